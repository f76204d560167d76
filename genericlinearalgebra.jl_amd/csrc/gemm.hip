// MFMA GEMM engine for gfx950: C = alpha*op(A)*op(B) + beta*C, FP64 on
// v_mfma_f64_16x16x4_f64 and FP32 on v_mfma_f32_16x16x4_f32.
//
// These are the dense contractions of the reference's hot path:
//   - rankUpdate!(Hermitian(C,:L), A, -1)   src/juliaBLAS.jl:89-112 (SYRK; mode 1)
//   - rdiv!(A21, LowerTriangular(A11)')     src/cholesky.jl:28,48  (TRSM via
//     the inverted diagonal block; NT form)
//   - mul!(M, V2', A2, 1, 1) / mul!(A2, V2, M, -1, 1)  src/householder.jl:144,149
//     (WY update; TN and NN forms)
//
// Design: 256 threads = 4 wave64s in a 2×2 arrangement; a workgroup tile is
// BM×BN and each wave owns (BM/2)×(BN/2) as a grid of 16×16 MFMA blocks with
// their accumulators in registers.  K advances BK=16 at a time: the next
// K-slab is fetched from HBM/L2 into registers while the current one feeds
// the MFMAs, then written to LDS — to the other of two buffers (one barrier
// per slab), or, for FP64 128×128 tiles by default, back into the single
// buffer after a second barrier (half the LDS).  A third form for FP64 NT
// products loads the slabs straight into an LDS ring (gemm_dl_kernel).
// LDS holds slabs K-major ([k][mn]) so that a
// fragment read — lane l takes element (mn = l%16, k = l/16) — is 16
// consecutive words per k row.  Tiles are mapped to workgroups XCD-aware:
// the dispatcher places workgroup b on XCD b%8, so consecutive tile ids are
// given to the same XCD to share its L2 for the A row-panels and B
// column-panels they have in common.
#include <algorithm>
#include <type_traits>

#include "gla_internal.hpp"

namespace gla {

namespace {

constexpr int kThreads = 256;
constexpr int kBK = 16;

// FP64 direct-to-LDS kernel selection (process-wide): 0 off, 1 (BK 8, ring
// of 4), 2 (BK 8, ring of 3; the default: the n = 16384 Cholesky's SYRK on
// 7/8 of the CUs runs 5% faster than with the register-staged kernel),
// 3 (BK 16, ring of 3).
int g_gemm_dl = 2;
// FP64 128×128 tiles with one LDS buffer (1, the default: measured 1–2%
// faster for the Cholesky, whose chain kernels then find LDS beside the
// trailing SYRK) or two (0).
int g_gemm_lds1 = 1;

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <typename T, int VEC>
struct VecT;
template <typename T>
struct VecT<T, 1> {
    typedef T type;
};
template <>
struct VecT<double, 2> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct VecT<float, 4> {
    typedef float type __attribute__((ext_vector_type(4)));
};

template <typename T>
struct GemmP {
    int64_t M, N, K;
    const T *A;
    int64_t lda;
    const T *B;
    int64_t ldb;
    T *C;
    int64_t ldc;
    T alpha, beta;
    int64_t diag_off;
    int tiles_m, tiles_n;
    int total;    // number of tiles actually enumerated
    int per_xcd;  // ceil(total / 8)
    int64_t kc;   // K extent of one split (split-K along blockIdx.y)
    T *ws;        // split-K partials: ws[z][M×N], or null for a direct epilogue
    int kfrom_row;  // 1: the K range of a tile starts at its first row (op(A)(i,k) = 0 for k < i)
    int batch;      // strided batch along blockIdx.z (1: a single product)
    int64_t sA, sB, sC;  // batch strides (elements)
};

// One operand slab: BX (M or N extent) × kBK, staged through registers into
// LDS as S[k * LDX + x].  MN_CONTIG: element (x,k) lives at P[x + k*ld]
// (column-major op(A) with 'N', or op(B) with 'T'); otherwise at P[k + x*ld].
template <typename T, int BX, bool MN_CONTIG, int VEC>
struct Slab {
    static constexpr int NV = BX * kBK / kThreads;
    static constexpr int NL = NV / VEC;
    static_assert(NV % VEC == 0 && NL >= 1, "slab shape");
    typedef typename VecT<T, VEC>::type vec_t;
    T r[NV];

    __device__ __forceinline__ void load(const T *__restrict__ P, int64_t ld, int64_t x0,
                                         int64_t k0, int64_t xlim, int64_t klim, bool full)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            int x, k;
            if (MN_CONTIG) {
                constexpr int TPC = BX / VEC, CPP = kThreads / TPC;
                x = (t % TPC) * VEC;
                k = t / TPC + l * CPP;
            } else {
                constexpr int TPX = kBK / VEC, XPP = kThreads / TPX;
                x = t / TPX + l * XPP;
                k = (t % TPX) * VEC;
            }
            const int64_t gx = x0 + x, gk = k0 + k;
            const T *src = MN_CONTIG ? P + gx + gk * ld : P + gk + gx * ld;
            if (full) {
                vec_t v = *reinterpret_cast<const vec_t *>(src);
#pragma unroll
                for (int e = 0; e < VEC; ++e) r[l * VEC + e] = reinterpret_cast<T *>(&v)[e];
            } else {
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    bool ok = MN_CONTIG ? (gx + e < xlim && gk < klim) : (gx < xlim && gk + e < klim);
                    r[l * VEC + e] = ok ? src[e] : T(0);
                }
            }
        }
    }

    template <int LDX>
    __device__ __forceinline__ void store(T *S) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            if (MN_CONTIG) {
                constexpr int TPC = BX / VEC, CPP = kThreads / TPC;
                const int x = (t % TPC) * VEC, k = t / TPC + l * CPP;
                vec_t v;
#pragma unroll
                for (int e = 0; e < VEC; ++e) reinterpret_cast<T *>(&v)[e] = r[l * VEC + e];
                *reinterpret_cast<vec_t *>(S + k * LDX + x) = v;
            } else {
                constexpr int TPX = kBK / VEC, XPP = kThreads / TPX;
                const int x = t / TPX + l * XPP, k = (t % TPX) * VEC;
#pragma unroll
                for (int e = 0; e < VEC; ++e) S[(k + e) * LDX + x] = r[l * VEC + e];
            }
        }
    }
};

// Row tiles [first, first+cnt) of tile column c that touch the written
// region: mode 1 keeps row + diag_off >= col, mode 2 keeps row + diag_off <= col.
template <int BM, int BN, int MODE>
__host__ __device__ __forceinline__ void tile_range(int c, int tiles_m, int64_t diag_off,
                                                    int &first, int &cnt)
{
    if (MODE == 1) {
        // first row tile whose last row reaches the diagonal at column c*BN
        int64_t need = (int64_t)c * BN - diag_off - (BM - 1);
        first = need <= 0 ? 0 : (int)((need + BM - 1) / BM);
        cnt = tiles_m - first;
    } else {
        // last row tile whose first row is on or above the diagonal at the
        // tile column's last column
        int64_t lim = (int64_t)c * BN + (BN - 1) - diag_off;
        first = 0;
        cnt = lim < 0 ? 0 : (int)(lim / BM) + 1;
        if (cnt > tiles_m) cnt = tiles_m;
    }
    if (cnt < 0) cnt = 0;
}

// Maps the linear (XCD-regrouped) tile id of this workgroup to (tm, tn).
// mode 1 enumerates only tiles that touch the "lower" region
// row + diag_off >= col, column by column.
template <int BM, int BN, int MODE, typename P>
__device__ __forceinline__ bool tile_of(const P &p, int &tm, int &tn)
{
    const int b = blockIdx.x;
    const int t = (b % 8) * p.per_xcd + b / 8;
    if (t >= p.total) return false;
    if (MODE == 0) {
        tm = t % p.tiles_m;
        tn = t / p.tiles_m;
        return true;
    }
    int rem = t;
    for (int c = 0; c < p.tiles_n; ++c) {
        int first, cnt;
        tile_range<BM, BN, MODE>(c, p.tiles_m, p.diag_off, first, cnt);
        if (rem < cnt) {
            tm = first + rem;
            tn = c;
            return true;
        }
        rem -= cnt;
    }
    return false;
}

// NBUF = 1: one LDS buffer (half the LDS: kernels of the caller's stream
// then fit beside two workgroups of this one on a CU) at the cost of a
// second barrier per K-slab.
template <typename T, int BM, int BN, bool TA, bool TB, int MODE, int VEC, int NBUF = 2>
__global__ __launch_bounds__(kThreads, 2) void gemm_kernel(GemmP<T> p)
{
    constexpr int PAD = 32 / sizeof(T);
    constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int MI = WM / 16, NJ = WN / 16;
    typedef typename Mfma<T>::acc_t acc_t;

    __shared__ __attribute__((aligned(16))) T As[NBUF][kBK * LDA_S];
    __shared__ __attribute__((aligned(16))) T Bs[NBUF][kBK * LDB_S];

    int tm, tn;
    if (!tile_of<BM, BN, MODE>(p, tm, tn)) return;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    // Split-K: this workgroup contracts K range [koff, koff + K).
    // Triangular operands (kfrom_row): terms with k below the tile's first
    // row are zero and skipped.
    const int64_t koff = (p.kfrom_row ? m0 : 0) + (int64_t)blockIdx.y * p.kc;
    int64_t K = p.K - koff < p.kc ? p.K - koff : p.kc;
    if (K < 0) K = 0;
    const int64_t bz = blockIdx.z;
    const T *Ap = p.A + bz * p.sA + (TA ? koff : koff * p.lda);
    const T *Bp = p.B + bz * p.sB + (TB ? koff * p.ldb : koff);

    // Accumulator element r of block (i,j) is C(row, col) with col = lane%16
    // inside the 16×16 block and row = 4*(lane/16) + r for the f32
    // instruction, row = lane/16 + 4*r for the f64 one (measured on gfx950:
    // the f64 16x16x4 result interleaves its four row groups).
    constexpr bool kF64Rows = sizeof(T) == 8;
    const bool split = p.ws != nullptr;
    // Accumulators start at zero; C is read in the epilogue.  (Starting them
    // at ±C instead stalls the first MFMA on the C loads: measured 15%
    // slower for K = 512 updates on gfx950.)
    acc_t acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // op(A) is M×K: MN-contiguous when not transposed.  op(B) is K×N:
    // MN-contiguous when transposed.
    Slab<T, BM, !TA, VEC> sa;
    Slab<T, BN, TB, VEC> sb;

    const int nk = (int)((K + kBK - 1) / kBK);
    const bool mfull = m0 + BM <= p.M, nfull = n0 + BN <= p.N;

    if (nk > 0) {
        bool kfull = kBK <= K;
        sa.load(Ap, p.lda, m0, 0, p.M, K, mfull && kfull);
        sb.load(Bp, p.ldb, n0, 0, p.N, K, nfull && kfull);
        sa.template store<LDA_S>(As[0]);
        sb.template store<LDB_S>(Bs[0]);
        __syncthreads();
    }

    for (int kt = 0; kt < nk; ++kt) {
        const int cur = NBUF == 2 ? (kt & 1) : 0;
        const bool more = kt + 1 < nk;
        if (more) {
            const int64_t k0 = (int64_t)(kt + 1) * kBK;
            bool kfull = k0 + kBK <= K;
            sa.load(Ap, p.lda, m0, k0, p.M, K, mfull && kfull);
            sb.load(Bp, p.ldb, n0, k0, p.N, K, nfull && kfull);
        }
        const T *as = As[cur];
        const T *bs = Bs[cur];
#pragma unroll
        for (int kk = 0; kk < kBK / 4; ++kk) {
            const int kr = kk * 4 + (lane >> 4);
            T a[MI], b[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) a[i] = as[kr * LDA_S + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NJ; ++j) b[j] = bs[kr * LDB_S + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
        if (more) {
            if (NBUF == 1) __syncthreads();   // every wave is done with the buffer
            sa.template store<LDA_S>(As[NBUF == 2 ? cur ^ 1 : 0]);
            sb.template store<LDB_S>(Bs[NBUF == 2 ? cur ^ 1 : 0]);
        }
        __syncthreads();
    }

    // Epilogue.
    const T alpha = split ? T(1) : p.alpha, beta = split ? T(0) : p.beta;
    const bool beta0 = beta == T(0);
    T *const Cp = split ? p.ws + (int64_t)blockIdx.y * p.M * p.N : p.C + bz * p.sC;
    const int64_t ldc = split ? p.M : p.ldc;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int64_t row = m0 + wm * WM + i * 16 + (kF64Rows ? (lane >> 4) : 4 * (lane >> 4));
            const int64_t col = n0 + wn * WN + j * 16 + (lane & 15);
            if (col >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gr = row + (kF64Rows ? 4 * r : r);
                T *c = Cp + gr + col * ldc;
                bool ok = gr < p.M;
                if (MODE == 1) ok = ok && (gr + p.diag_off >= col);
                if (MODE == 2) ok = ok && (gr + p.diag_off <= col);
                if (ok) {
                    T v = alpha * acc[i][j][r];
                    if (!beta0) v += beta * *c;
                    *c = v;
                }
            }
        }
}

// Direct-to-LDS variant for FP64 with MN-contiguous operands (op(A) = A
// 'N', op(B) = Bᵀ 'T': the SYRK and NT products), 128×128 tiles.  Each
// K-row of a slab (128 doubles, 1 KB) is one global_load_lds_dwordx4 of one
// wave, written straight into an NS-deep LDS ring without passing through
// registers; the waves wait only for the slab they are about to use
// (vmcnt), so NS − 1 slabs stay in flight behind the MFMAs.  Rows past M or
// N load clamped (valid) addresses whose results are never written; K must
// be a multiple of BKD.
template <int BKD, int NS, int MODE>
__global__ __launch_bounds__(kThreads, (2 * NS * BKD * 132 * 8 <= 81920 ? 2 : 1)) void
gemm_dl_kernel(GemmP<double> p)
{
    typedef double T;
    constexpr int BM = 128, BN = 128, LDX = 132;
    constexpr int WM = 64, WN = 64, MI = 4, NJ = 4;
    constexpr int LPW = BKD / 4;   // K-rows of each operand a wave loads per slab
    static_assert(BKD % 4 == 0 && NS >= 3 && NS <= 4, "slab shape");
    typedef Mfma<T>::acc_t acc_t;
    __shared__ __attribute__((aligned(16))) T As[NS][BKD * LDX];
    __shared__ __attribute__((aligned(16))) T Bs[NS][BKD * LDX];

    int tm, tn;
    if (!tile_of<BM, BN, MODE>(p, tm, tn)) return;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t koff = (p.kfrom_row ? m0 : 0);
    int64_t K = p.K - koff;
    if (K < 0) K = 0;
    const int64_t bz = blockIdx.z;
    const T *Ap = p.A + bz * p.sA + koff * p.lda;
    const T *Bp = p.B + bz * p.sB + koff * p.ldb;
    const int64_t xa = m0 + 2 * lane < p.M - 1 ? m0 + 2 * lane : p.M - 2;
    const int64_t xb = n0 + 2 * lane < p.N - 1 ? n0 + 2 * lane : p.N - 2;

    acc_t acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nk = (int)(K / BKD);
    auto issue = [&](int sl) {
        const int st = sl % NS;
        const int64_t k0 = (int64_t)sl * BKD;
#pragma unroll
        for (int r = 0; r < LPW; ++r) {
            const int kr = wave * LPW + r;
            __builtin_amdgcn_global_load_lds(
                (const void *)(Ap + xa + (k0 + kr) * p.lda),
                (__attribute__((address_space(3))) void *)(&As[st][kr * LDX]), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(
                (const void *)(Bp + xb + (k0 + kr) * p.ldb),
                (__attribute__((address_space(3))) void *)(&Bs[st][kr * LDX]), 16, 0, 0);
        }
    };
    // vmcnt(n): n of this wave's vector loads may still be outstanding
    constexpr int PER = 2 * LPW;   // loads per wave per slab
    auto wait_slab = [&](int later) {   // later = slabs issued after the one needed
        if (NS == 4 && later >= 2)
            __builtin_amdgcn_s_waitcnt(((2 * PER) & 15) | (((2 * PER) >> 4) << 14) | 0x0F70);
        else if (later >= 1)
            __builtin_amdgcn_s_waitcnt((PER & 15) | ((PER >> 4) << 14) | 0x0F70);
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);
    };
    for (int sl = 0; sl < NS - 1 && sl < nk; ++sl) issue(sl);
    for (int sl = 0; sl < nk; ++sl) {
        wait_slab(min(NS - 2, nk - 1 - sl));
        __syncthreads();
        if (sl + NS - 1 < nk) issue(sl + NS - 1);
        const T *as = As[sl % NS];
        const T *bs = Bs[sl % NS];
#pragma unroll
        for (int kk = 0; kk < BKD / 4; ++kk) {
            const int kr = kk * 4 + (lane >> 4);
            T a[MI], b[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) a[i] = as[kr * LDX + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NJ; ++j) b[j] = bs[kr * LDX + wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
        }
    }

    const T alpha = p.alpha, beta = p.beta;
    const bool beta0 = beta == T(0);
    T *const Cp = p.C + bz * p.sC;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int64_t row = m0 + wm * WM + i * 16 + (lane >> 4);
            const int64_t col = n0 + wn * WN + j * 16 + (lane & 15);
            if (col >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gr = row + 4 * r;
                T *c = Cp + gr + col * p.ldc;
                bool ok = gr < p.M;
                if (MODE == 1) ok = ok && (gr + p.diag_off >= col);
                if (MODE == 2) ok = ok && (gr + p.diag_off <= col);
                if (ok) {
                    T v = alpha * acc[i][j][r];
                    if (!beta0) v += beta * *c;
                    *c = v;
                }
            }
        }
}

// C = alpha * Σ_z ws[z] + beta * C over the written region (split-K tail).
template <typename T, int MODE>
__global__ void splitk_reduce_kernel(GemmP<T> p, int splits)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= p.M) return;
    if (MODE == 1 && i + p.diag_off < j) return;
    if (MODE == 2 && i + p.diag_off > j) return;
    const int64_t mn = p.M * p.N;
    T acc = T(0);
    for (int z = 0; z < splits; ++z) acc += p.ws[z * mn + i + j * p.M];
    T *c = p.C + i + j * p.ldc;
    T v = p.alpha * acc;
    if (p.beta != T(0)) v += p.beta * *c;
    *c = v;
}

template <typename T, int BM, int BN, bool TA, bool TB, int MODE, int VEC>
int launch(hipStream_t s, GemmP<T> p, int splits)
{
    p.tiles_m = (int)cdiv(p.M, BM);
    p.tiles_n = (int)cdiv(p.N, BN);
    if (MODE == 0) {
        p.total = p.tiles_m * p.tiles_n;
    } else {
        int64_t tot = 0;
        for (int c = 0; c < p.tiles_n; ++c) {
            int first, cnt;
            tile_range<BM, BN, MODE>(c, p.tiles_m, p.diag_off, first, cnt);
            tot += cnt;
        }
        p.total = (int)tot;
    }
    if (p.total == 0) return GLA_OK;
    p.per_xcd = (int)cdiv(p.total, 8);
    if (splits <= 1) {
        p.ws = nullptr;
        p.kc = p.K > 0 ? p.K : 1;
        splits = 1;
    } else {
        p.kc = cdiv(cdiv(p.K, splits), kBK) * kBK;
        splits = (int)cdiv(p.K, p.kc);
    }
    if (p.batch > 1) {
        p.ws = nullptr;   // split-K partials are single-product scratch
        p.kc = p.K > 0 ? p.K : 1;
        splits = 1;
    }
    if constexpr (std::is_same<T, double>::value && BM == 128 && BN == 128 && !TA && TB &&
                  VEC == 2) {
        const int dl = g_gemm_dl;
        const int bkd = dl == 3 ? 16 : 8;
        if (dl > 0 && splits == 1 && p.M >= 2 && p.N >= 2 && p.K % bkd == 0) {
            const dim3 grid(8 * p.per_xcd, 1, p.batch > 1 ? p.batch : 1);
            if (dl == 1)
                hipLaunchKernelGGL((gemm_dl_kernel<8, 4, MODE>), grid, dim3(kThreads), 0, s, p);
            else if (dl == 2)
                hipLaunchKernelGGL((gemm_dl_kernel<8, 3, MODE>), grid, dim3(kThreads), 0, s, p);
            else
                hipLaunchKernelGGL((gemm_dl_kernel<16, 3, MODE>), grid, dim3(kThreads), 0, s, p);
            GLA_LAUNCH_CHECK();
            return GLA_OK;
        }
    }
    if constexpr (std::is_same<T, double>::value && BM == 128 && BN == 128) {
        if (g_gemm_lds1) {
            hipLaunchKernelGGL((gemm_kernel<T, BM, BN, TA, TB, MODE, VEC, 1>),
                               dim3(8 * p.per_xcd, splits, p.batch > 1 ? p.batch : 1),
                               dim3(kThreads), 0, s, p);
            GLA_LAUNCH_CHECK();
            if (p.ws) {
                hipLaunchKernelGGL((splitk_reduce_kernel<T, MODE>),
                                   dim3((unsigned)cdiv(p.M, 256), p.N), dim3(256), 0, s, p, splits);
                GLA_LAUNCH_CHECK();
            }
            return GLA_OK;
        }
    }
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, TA, TB, MODE, VEC>),
                       dim3(8 * p.per_xcd, splits, p.batch > 1 ? p.batch : 1), dim3(kThreads), 0,
                       s, p);
    GLA_LAUNCH_CHECK();
    if (p.ws) {
        hipLaunchKernelGGL((splitk_reduce_kernel<T, MODE>), dim3((unsigned)cdiv(p.M, 256), p.N),
                           dim3(256), 0, s, p, splits);
        GLA_LAUNCH_CHECK();
    }
    return GLA_OK;
}

template <typename T, bool TA, bool TB, int MODE, int VEC>
int pick_tile(hipStream_t s, const GemmP<T> &p, size_t ws_elems)
{
    // 128×128 tiles when the problem yields enough of them to fill the
    // chip; 64×64 otherwise, split along K when even those are too few and
    // K is long (tall-skinny products such as Vᵀ A in the WY update).
    int64_t big = cdiv(p.M, 128) * cdiv(p.N, 128);
    if (MODE != 0) big /= 2;
    if (big >= 384) return launch<T, 128, 128, TA, TB, MODE, VEC>(s, p, 1);
    int64_t small = cdiv(p.M, 64) * cdiv(p.N, 64);
    if (MODE != 0) small = (small + 1) / 2;
    int splits = 1;
    if (p.ws && small < 256 && p.K >= 1024) {
        splits = (int)std::min<int64_t>(cdiv(p.K, 512), cdiv(512, small));
        const int64_t mn = p.M * p.N;
        if (mn > 0) splits = (int)std::min<int64_t>(splits, (int64_t)(ws_elems / (size_t)mn));
    }
    return launch<T, 64, 64, TA, TB, MODE, VEC>(s, p, splits);
}

template <typename T, int VEC>
int dispatch(hipStream_t s, char ta, char tb, int mode, const GemmP<T> &p, size_t ws_elems)
{
    if (mode == 0) {
        if (ta == 'N' && tb == 'N') return pick_tile<T, false, false, 0, VEC>(s, p, ws_elems);
        if (ta == 'N' && tb == 'T') return pick_tile<T, false, true, 0, VEC>(s, p, ws_elems);
        if (ta == 'T' && tb == 'N') return pick_tile<T, true, false, 0, VEC>(s, p, ws_elems);
    } else if (mode == 1) {
        if (ta == 'N' && tb == 'T') return pick_tile<T, false, true, 1, VEC>(s, p, ws_elems);
    } else if (mode == 2) {
        if (ta == 'N' && tb == 'T') return pick_tile<T, false, true, 2, VEC>(s, p, ws_elems);
    }
    set_error("gemm: unsupported combination transa=%c transb=%c mode=%d", ta, tb, mode);
    return GLA_ERR_UNSUPPORTED;
}

}  // namespace

template <typename T>
int gemm(hipStream_t s, char ta, char tb, int mode, int64_t M, int64_t N, int64_t K, T alpha,
         const T *A, int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc,
         int64_t diag_off, T *ws, size_t ws_elems, int kfrom_row)
{
    if (M <= 0 || N <= 0) return GLA_OK;
    GemmP<T> p{};
    p.kfrom_row = kfrom_row;
    if (kfrom_row) ws_elems = 0;
    p.M = M;
    p.N = N;
    p.K = K < 0 ? 0 : K;
    p.A = A;
    p.lda = lda;
    p.B = B;
    p.ldb = ldb;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    p.diag_off = diag_off;
    p.ws = ws_elems > 0 ? ws : nullptr;
    p.batch = 1;
    constexpr int VW = 16 / sizeof(T);
    const bool aligned = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) &&
                         (lda % VW == 0) && (ldb % VW == 0);
    if (aligned) return dispatch<T, VW>(s, ta, tb, mode, p, ws_elems);
    return dispatch<T, 1>(s, ta, tb, mode, p, ws_elems);
}

void set_gemm_dl(int v) { g_gemm_dl = v; }
void set_gemm_lds1(int v) { g_gemm_lds1 = v; }
int get_gemm_dl() { return g_gemm_dl; }
int get_gemm_lds1() { return g_gemm_lds1; }

template <typename T>
int gemm_batched(hipStream_t s, char ta, char tb, int mode, int64_t M, int64_t N, int64_t K,
                 T alpha, const T *A, int64_t lda, int64_t sA, const T *B, int64_t ldb,
                 int64_t sB, T beta, T *C, int64_t ldc, int64_t sC, int batch, int kfrom_row)
{
    if (M <= 0 || N <= 0 || batch <= 0) return GLA_OK;
    GemmP<T> p{};
    p.kfrom_row = kfrom_row;
    p.M = M;
    p.N = N;
    p.K = K < 0 ? 0 : K;
    p.A = A;
    p.lda = lda;
    p.B = B;
    p.ldb = ldb;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    p.batch = batch;
    p.sA = sA;
    p.sB = sB;
    p.sC = sC;
    constexpr int VW = 16 / sizeof(T);
    const bool aligned = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) &&
                         (lda % VW == 0) && (ldb % VW == 0) && (sA % VW == 0) && (sB % VW == 0);
    if (aligned) return dispatch<T, VW>(s, ta, tb, mode, p, 0);
    return dispatch<T, 1>(s, ta, tb, mode, p, 0);
}

template int gemm<double>(hipStream_t, char, char, int, int64_t, int64_t, int64_t, double,
                          const double *, int64_t, const double *, int64_t, double, double *,
                          int64_t, int64_t, double *, size_t, int);
template int gemm<float>(hipStream_t, char, char, int, int64_t, int64_t, int64_t, float,
                         const float *, int64_t, const float *, int64_t, float, float *,
                         int64_t, int64_t, float *, size_t, int);
template int gemm_batched<double>(hipStream_t, char, char, int, int64_t, int64_t, int64_t, double,
                                  const double *, int64_t, int64_t, const double *, int64_t,
                                  int64_t, double, double *, int64_t, int64_t, int, int);
template int gemm_batched<float>(hipStream_t, char, char, int, int64_t, int64_t, int64_t, float,
                                 const float *, int64_t, int64_t, const float *, int64_t, int64_t,
                                 float, float *, int64_t, int64_t, int, int);

}  // namespace gla
