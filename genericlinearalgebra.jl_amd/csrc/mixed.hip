// Mixed-precision Cholesky solve with FP64 iterative refinement.
// No counterpart in the reference (its Cholesky runs in one precision,
// src/cholesky.jl).  A x = B for SPD A (n×n, FP64) and nrhs right-hand sides:
//
//   factor   L Lᵀ ≈ A in a working copy W:
//     'd'  FP64 throughout (the plain FP64 path, cholRecursive! on device);
//     's'  FP32 throughout;
//     'b'/'h'  FP32 storage, every 128×128 diagonal block factored in FP64,
//          the panel solve in FP32, and the trailing SYRK/GEMM updates
//          (rankUpdate!, src/juliaBLAS.jl:89) on BF16 / FP16 MFMA with FP32
//          accumulation (v_mfma_f32_16x16x32_{bf16,f16});
//   refine   r = B − A x in FP64 (MFMA GEMM), x += (L Lᵀ)⁻¹ r with the
//            low-precision factor, until the normwise backward error
//            η = max_k ‖r_k‖ / (‖A‖_F ‖x_k‖ + ‖b_k‖) ≤ tol.
//
// The triangular solves reuse the right-side solve B <- B L⁻ᵀ of the
// Cholesky driver on the transposed right-hand sides: L y = r is
// yᵀ = rᵀ L⁻ᵀ; Lᵀ z = y is a lower solve with the reversed matrix
// L' = P Lᵀ P (P the reversal), z = P (L'⁻¹ (P y)).
#include <algorithm>
#include <cmath>
#include <vector>

#include "gla_internal.hpp"

namespace gla {

namespace {

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

// ---------------------------------------------------- low-precision GEMM --
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <typename TIN>
struct LpMfma;
template <>
struct LpMfma<__bf16> {
    typedef bf16x8 v8;
    static __device__ __forceinline__ f4 mma(v8 a, v8 b, f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct LpMfma<_Float16> {
    typedef f16x8 v8;
    static __device__ __forceinline__ f4 mma(v8 a, v8 b, f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

struct LpP {
    int64_t M, N, K;
    const float *A;
    int64_t lda;
    const float *B;
    int64_t ldb;
    float *C;
    int64_t ldc;
    float alpha;   // C <- C + alpha A Bᵀ
    int mode;      // 0 all, 1 lower (row >= col)
    int tiles_m, tiles_n, total, per_xcd;
};

constexpr int kLpT = 256, kLpBM = 128, kLpBK = 32, kLpLD = kLpBK + 8;

__device__ __forceinline__ void lp_tile_range(int c, int tiles_m, int mode, int &first, int &cnt)
{
    if (mode == 0) {
        first = 0;
        cnt = tiles_m;
        return;
    }
    const int64_t need = (int64_t)c * kLpBM - (kLpBM - 1);
    first = need <= 0 ? 0 : (int)((need + kLpBM - 1) / kLpBM);
    cnt = tiles_m - first;
    if (cnt < 0) cnt = 0;
}

// C (M×N, FP32) += alpha · A Bᵀ with A (M×K) and B (N×K) column-major FP32,
// both rounded to TIN as they are staged into LDS; FP32 accumulation.
template <typename TIN>
__global__ __launch_bounds__(kLpT, 2) void lp_gemm_nt_kernel(LpP p)
{
    typedef typename LpMfma<TIN>::v8 v8;
    __shared__ __attribute__((aligned(16))) TIN As[2][kLpBM * kLpLD];
    __shared__ __attribute__((aligned(16))) TIN Bs[2][kLpBM * kLpLD];

    const int b = blockIdx.x;
    const int tid = (b % 8) * p.per_xcd + b / 8;
    if (tid >= p.total) return;
    int tm = 0, tn = 0;
    {
        int rem = tid;
        for (int c = 0; c < p.tiles_n; ++c) {
            int first, cnt;
            lp_tile_range(c, p.tiles_m, p.mode, first, cnt);
            if (rem < cnt) {
                tm = first + rem;
                tn = c;
                break;
            }
            rem -= cnt;
        }
    }
    const int64_t m0 = (int64_t)tm * kLpBM, n0 = (int64_t)tn * kLpBM;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    // staging: thread t holds row m = t % 128 of the slab, k = 16·(t / 128)
    // .. +15: each load instruction reads 64 consecutive rows (coalesced);
    // the 16 values are rounded to TIN and written as two 16-byte vectors in
    // the [m][k] layout the MFMA operands are read from.
    float ra[16], rb[16];
    const int sm = t & 127, sk = 16 * (t >> 7);
    auto gload = [&](const float *X, int64_t ldx, int64_t x0, int64_t xlim, int64_t k0, float *r) {
        const int64_t gx = x0 + sm;
        const bool okx = gx < xlim;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int64_t gk = k0 + sk + e;
            r[e] = (okx && gk < p.K) ? X[gx + gk * ldx] : 0.f;
        }
    };
    auto lstore = [&](TIN *S, const float *r) {
        v8 lo, hi;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            lo[e] = (TIN)r[e];
            hi[e] = (TIN)r[8 + e];
        }
        *reinterpret_cast<v8 *>(&S[sm * kLpLD + sk]) = lo;
        *reinterpret_cast<v8 *>(&S[sm * kLpLD + sk + 8]) = hi;
    };

    const int nk = (int)((p.K + kLpBK - 1) / kLpBK);
    if (nk > 0) {
        gload(p.A, p.lda, m0, p.M, 0, ra);
        gload(p.B, p.ldb, n0, p.N, 0, rb);
        lstore(As[0], ra);
        lstore(Bs[0], rb);
        __syncthreads();
    }
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            gload(p.A, p.lda, m0, p.M, (int64_t)(kt + 1) * kLpBK, ra);
            gload(p.B, p.ldb, n0, p.N, (int64_t)(kt + 1) * kLpBK, rb);
        }
        v8 a[4], bb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = *reinterpret_cast<const v8 *>(&As[cur][(wm * 64 + i * 16 + (lane & 15)) * kLpLD +
                                                          8 * (lane >> 4)]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bb[j] = *reinterpret_cast<const v8 *>(&Bs[cur][(wn * 64 + j * 16 + (lane & 15)) * kLpLD +
                                                           8 * (lane >> 4)]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = LpMfma<TIN>::mma(a[i], bb[j], acc[i][j]);
        if (more) {
            lstore(As[cur ^ 1], ra);
            lstore(Bs[cur ^ 1], rb);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4);
            const int64_t col = n0 + wn * 64 + j * 16 + (lane & 15);
            if (col >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gr = row + r;
                if (gr < p.M && (p.mode == 0 || gr >= col)) {
                    float *c = p.C + gr + col * p.ldc;
                    *c = *c + p.alpha * acc[i][j][r];
                }
            }
        }
}

template <typename TIN>
int lp_gemm_nt(hipStream_t s, int mode, int64_t M, int64_t N, int64_t K, float alpha,
               const float *A, int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc)
{
    if (M <= 0 || N <= 0) return GLA_OK;
    LpP p{};
    p.M = M;
    p.N = N;
    p.K = K;
    p.A = A;
    p.lda = lda;
    p.B = B;
    p.ldb = ldb;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.mode = mode;
    p.tiles_m = (int)cdiv(M, kLpBM);
    p.tiles_n = (int)cdiv(N, kLpBM);
    int64_t tot = 0;
    for (int c = 0; c < p.tiles_n; ++c) {
        int first, cnt;
        if (mode == 0) {
            cnt = p.tiles_m;
        } else {
            const int64_t need = (int64_t)c * kLpBM - (kLpBM - 1);
            first = need <= 0 ? 0 : (int)((need + kLpBM - 1) / kLpBM);
            cnt = std::max(0, p.tiles_m - first);
        }
        tot += cnt;
    }
    p.total = (int)tot;
    if (p.total == 0) return GLA_OK;
    p.per_xcd = (int)cdiv(p.total, 8);
    hipLaunchKernelGGL((lp_gemm_nt_kernel<TIN>), dim3(8 * p.per_xcd), dim3(kLpT), 0, s, p);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// ------------------------------------------------------------- helpers --
// Element-wise kernels over an m×ncols index space: x covers rows, y strides
// over columns (the grid's y extent is capped at 65535).
#define GLA_COLS(j, ncols) for (int64_t j = blockIdx.y; j < (ncols); j += gridDim.y)
#define GLA_ROW(i) const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x

// dst (FP32, lower triangle) <- src (FP64, lower triangle), n×n
__global__ void cvt_lower_d2s(int64_t n, const double *src, int64_t lds, float *dst, int64_t ldd)
{
    GLA_ROW(i);
    GLA_COLS(j, n) if (i < n && i >= j) dst[i + j * ldd] = (float)src[i + j * lds];
}
template <typename TD, typename TS>
__global__ void cvt_block(int nb, const TS *src, int64_t lds, TD *dst, int64_t ldd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < nb && i >= j) dst[i + (int64_t)j * ldd] = (TD)src[i + (int64_t)j * lds];
}
template <typename TD, typename TS>
__global__ void cvt_vec(int64_t n, const TS *src, TD *dst)
{
    GLA_ROW(i);
    if (i < n) dst[i] = (TD)src[i];
}
// Lr[i][j] = L[n-1-j][n-1-i] for i >= j: the reversal P Lᵀ P, lower.
template <typename T>
__global__ void reverse_transpose_lower(int64_t n, const T *L, int64_t ldl, T *Lr, int64_t ldr)
{
    GLA_ROW(i);
    GLA_COLS(j, n) if (i < n && i >= j) Lr[i + j * ldr] = L[(n - 1 - j) + (n - 1 - i) * ldl];
}
__global__ void copy2d_d(int64_t m, int64_t n, const double *S, int64_t lds, double *D, int64_t ldd)
{
    GLA_ROW(i);
    GLA_COLS(j, n) if (i < m) D[i + j * ldd] = S[i + j * lds];
}
__global__ void zero2d_d(int64_t m, int64_t n, double *D, int64_t ldd)
{
    GLA_ROW(i);
    GLA_COLS(j, n) if (i < m) D[i + j * ldd] = 0.0;
}
#undef GLA_COLS
#undef GLA_ROW

inline dim3 g2(int64_t rows, int64_t cols)
{
    return dim3((unsigned)cdiv(rows, 256), (unsigned)std::min<int64_t>(cols, 65535));
}

// out[blockIdx.y] = Σ of squares over the lower triangle weighted to the
// full symmetric matrix (off-diagonal counted twice): ‖A‖_F² by columns.
__global__ __launch_bounds__(256) void sym_sumsq_cols(int64_t n, const double *A, int64_t lda,
                                                      double *out)
{
    __shared__ double red[256];
    const int64_t j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = j + threadIdx.x; i < n; i += 256) {
        const double v = A[i + j * lda];
        s += (i == j ? 1.0 : 2.0) * v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}


// ------------------------------------------------------- mixed factor --
// Blocked right-looking Cholesky of W (FP32, lower) with FP64 diagonal
// blocks and low-precision trailing updates, depth-1 look-ahead.
template <typename TIN>
int chol_mixed(Handle *h, hipStream_t s, int64_t n, float *W, int64_t ldw, double *D64,
               double *Linv64, float *Linv32, bool use_side = true)
{
    const int nb = 128;
    const int64_t nsteps = cdiv(n, nb);
    const bool overlap = use_side && nsteps > 2;
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
        GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += nb) {
        const int kb = (int)std::min<int64_t>(nb, n - k);
        const int64_t below = n - k - kb;
        float *Wkk = W + k + k * ldw;
        float *W21 = Wkk + kb;
        // diagonal block in FP64
        hipLaunchKernelGGL((cvt_block<double, float>), dim3(1, kb), dim3(128), 0, s, kb, Wkk, ldw,
                           D64, (int64_t)nb);
        GLA_LAUNCH_CHECK();
        GLA_TRY(chol_factor_block<double>(s, kb, D64, nb, h->dinfo, k, Linv64));
        hipLaunchKernelGGL((cvt_block<float, double>), dim3(1, kb), dim3(128), 0, s, kb, D64,
                           (int64_t)nb, Wkk, ldw);
        GLA_LAUNCH_CHECK();
        hipLaunchKernelGGL((cvt_vec<float, double>), dim3(8), dim3(256), 0, s, (int64_t)8 * 256,
                           Linv64, Linv32);
        GLA_LAUNCH_CHECK();
        GLA_TRY(chol_solve_block<float>(s, kb, Wkk, ldw, Linv32, W21, ldw, below));
        if (below == 0) break;
        const int kb2 = (int)std::min<int64_t>(nb, below);
        const int64_t rest = below - kb2;
        if (overlap && rest > 0) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        GLA_TRY(lp_gemm_nt<TIN>(s, 1, below, kb2, kb, -1.f, W21, ldw, W21, ldw, W21 + kb * ldw, ldw));
        if (rest > 0) {
            hipStream_t us = overlap ? h->side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_step[0], 0));
            float *X2 = W21 + kb2;
            GLA_TRY(lp_gemm_nt<TIN>(us, 1, rest, rest, kb, -1.f, X2, ldw, X2, ldw,
                                    X2 + (kb + kb2) * ldw, ldw));
            if (overlap) GLA_HIP_TRY(hipEventRecord(h->ev_step[1], h->side));
        }
    }
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, h->side));
        GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    return GLA_OK;
}

// Wide-block form (NB = 512): the NB×NB diagonal block by the 128-wide
// driver above on the caller's stream, the panel below it by the FP32
// recursive solve, then the look-ahead trailing updates of cholBlocked!'s
// loop (src/cholesky.jl:17-35) with K = NB on the low-precision MFMA GEMM:
// the next block column on the caller's stream, the rest on the side stream.
template <typename TIN>
int chol_mixed_la(Handle *h, hipStream_t s, int64_t n, float *W, int64_t ldw, double *D64,
                  double *Linv64, float *Linv32, float *Linv, int NB)
{
    if (n <= NB) return chol_mixed<TIN>(h, s, n, W, ldw, D64, Linv64, Linv32, false);
    const int64_t nsteps = cdiv(n, NB);
    const bool overlap = nsteps > 2;
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_fork, s));
        GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    }
    int64_t k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += NB) {
        const int64_t kb = std::min<int64_t>(NB, n - k);
        const int64_t below = n - k - kb;
        float *Wkk = W + k + k * ldw;
        float *W21 = Wkk + kb;
        GLA_TRY(chol_mixed<TIN>(h, s, kb, Wkk, ldw, D64, Linv64, Linv32, false));
        if (below == 0) break;
        GLA_TRY(tile_inverses<float>(s, kb, Wkk, ldw, Linv));
        GLA_TRY(trsm_rlt<float>(s, below, kb, Wkk, ldw, Linv, W21, ldw));
        const int64_t kb2 = std::min<int64_t>(NB, below);
        const int64_t rest = below - kb2;
        if (overlap && rest > 0) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        GLA_TRY(lp_gemm_nt<TIN>(s, 1, below, kb2, kb, -1.f, W21, ldw, W21, ldw, W21 + kb * ldw, ldw));
        if (rest > 0) {
            hipStream_t us = overlap ? h->side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_step[0], 0));
            float *X2 = W21 + kb2;
            GLA_TRY(lp_gemm_nt<TIN>(us, 1, rest, rest, kb, -1.f, X2, ldw, X2, ldw,
                                    X2 + (kb + kb2) * ldw, ldw));
            if (overlap) GLA_HIP_TRY(hipEventRecord(h->ev_step[1], h->side));
        }
    }
    if (overlap) {
        GLA_HIP_TRY(hipEventRecord(h->ev_join, h->side));
        GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    return GLA_OK;
}

__global__ void fill_int(int64_t n, int *p, int v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_dbl(int64_t n, double *p)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// η = max_k ‖r_k‖ / (‖A‖ ‖x_k‖ + ‖b_k‖); a NaN in any column makes η NaN.
double backward_error(const double *rr, const double *xx, const double *bb, int64_t nrhs,
                      double anorm)
{
    double eta = 0.0;
    for (int64_t k = 0; k < nrhs; ++k) {
        const double den = anorm * std::sqrt(xx[k]) + std::sqrt(bb[k]);
        const double e = den > 0 ? std::sqrt(rr[k]) / den : std::sqrt(rr[k]);
        if (!(e <= eta)) eta = e;
    }
    return eta;
}

// Workspace of one mixed solve (element type T of the factor).
template <typename T>
struct MixedWS {
    T *W, *Lr, *Linv, *Linvr, *Dt, *Dtr, *Rt, *Yt, *Zt;
    double *Xt, *nrm, *D64, *Linv64;
    float *Linv32;
    int *flags, *err;
    int64_t nblk, npad, nchunk;

    static size_t bytes(int64_t n, int64_t nrhs)
    {
        const int64_t nblk = cdiv(n, 128), npad = nblk * 128, nch = cdiv(nrhs, 16);
        const size_t es = sizeof(T), r = 255;
        auto al = [&](size_t b) { return (b + r) / 256 * 256; };
        return 2 * al(es * (size_t)n * n) + 2 * al(es * tile_inverse_elems(n)) +
               2 * al(es * (size_t)npad * 128) + al(es * (size_t)npad * 16 * nch) +
               2 * al(es * (size_t)npad * 16) + al(8 * symv_scratch(n)) +
               al(8 * (size_t)(32 * nch + n + 64)) + al(8 * 128 * 128) + al(8 * 8 * 256) +
               al(4 * 8 * 256) + al(4 * (size_t)(nblk + 64)) + 4096;
    }
    void carve(void *base, int64_t n, int64_t nrhs)
    {
        nblk = cdiv(n, 128);
        npad = nblk * 128;
        nchunk = cdiv(nrhs, 16);
        char *p = reinterpret_cast<char *>(base);
        auto take = [&](size_t b) {
            char *q = p;
            p += (b + 255) / 256 * 256;
            return q;
        };
        const size_t es = sizeof(T);
        W = reinterpret_cast<T *>(take(es * (size_t)n * n));
        Lr = reinterpret_cast<T *>(take(es * (size_t)n * n));
        Linv = reinterpret_cast<T *>(take(es * tile_inverse_elems(n)));
        Linvr = reinterpret_cast<T *>(take(es * tile_inverse_elems(n)));
        Dt = reinterpret_cast<T *>(take(es * (size_t)npad * 128));
        Dtr = reinterpret_cast<T *>(take(es * (size_t)npad * 128));
        Rt = reinterpret_cast<T *>(take(es * (size_t)npad * 16 * nchunk));
        Yt = reinterpret_cast<T *>(take(es * (size_t)npad * 16));
        Zt = reinterpret_cast<T *>(take(es * (size_t)npad * 16));
        Xt = reinterpret_cast<double *>(take(8 * symv_scratch(n)));
        nrm = reinterpret_cast<double *>(take(8 * (size_t)(32 * nchunk + n + 64)));
        D64 = reinterpret_cast<double *>(take(8 * 128 * 128));
        Linv64 = reinterpret_cast<double *>(take(8 * 8 * 256));
        Linv32 = reinterpret_cast<float *>(take(4 * 8 * 256));
        flags = reinterpret_cast<int *>(take(4 * (size_t)(nblk + 64)));
        err = flags + nblk;
    }
};

// Solve structures of the factor in ws.W: tile inverses, diagonal-block
// inverses, and the same for the reversed Lr = J Lᵀ J (backward solve).
template <typename T>
int prepare_solves(hipStream_t s, int64_t n, MixedWS<T> &ws)
{
    GLA_TRY(tile_inverses<T>(s, n, ws.W, n, ws.Linv));
    GLA_TRY(diag_block_inverses<T>(s, n, ws.W, n, ws.Linv, ws.Dt));
    hipLaunchKernelGGL((reverse_transpose_lower<T>), g2(n, n), dim3(256), 0, s, n, ws.W, n, ws.Lr,
                       n);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses<T>(s, n, ws.Lr, n, ws.Linvr));
    return diag_block_inverses<T>(s, n, ws.Lr, n, ws.Linvr, ws.Dtr);
}

// Refinement outcome.
enum { kConverged = 0, kMaxIter = 1, kStalled = 2 };

// X = (L Lᵀ)⁻¹ B refined in FP64 until η ≤ tol.  Stops as stalled when η
// fails to halve over two corrections or is not finite.
template <typename T>
int refine(Handle *h, hipStream_t s, int64_t n, const double *A, int64_t lda, int64_t nrhs,
           const double *B, int64_t ldb, double *X, int64_t ldx, MixedWS<T> &ws, double anorm,
           double tol, int max_iter, int &epoch, int *iters, double *berr, int *outcome)
{
    const int64_t nch = ws.nchunk;
    std::vector<double> hn((size_t)32 * nch), bb((size_t)nrhs), rr((size_t)nrhs), xx((size_t)nrhs);
    hipLaunchKernelGGL(zero2d_d, g2(n, nrhs), dim3(256), 0, s, n, nrhs, X, ldx);
    GLA_LAUNCH_CHECK();
    double eta = INFINITY, prev2 = INFINITY, prev1 = INFINITY;
    int corr = 0;
    *outcome = kMaxIter;
    for (int it = 0;; ++it) {
        // r = b − A x per chunk of 16 columns, with ‖r_k‖², ‖x_k‖²
        hipLaunchKernelGGL(fill_dbl, dim3((unsigned)cdiv(32 * nch, 256)), dim3(256), 0, s, 32 * nch,
                           ws.nrm);
        GLA_LAUNCH_CHECK();
        for (int64_t c = 0; c < nch; ++c) {
            const int nr = (int)std::min<int64_t>(16, nrhs - 16 * c);
            GLA_TRY(symv_residual<T>(s, n, A, lda, X + 16 * c * ldx, ldx, B + 16 * c * ldb, ldb, nr,
                                     ws.Xt, ws.Rt + c * ws.npad * 16, ws.nrm + 32 * c));
        }
        int herr = 0;
        GLA_HIP_TRY(hipMemcpyAsync(hn.data(), ws.nrm, sizeof(double) * 32 * nch,
                                   hipMemcpyDeviceToHost, s));
        GLA_HIP_TRY(hipMemcpyAsync(&herr, ws.err, sizeof(int), hipMemcpyDeviceToHost, s));
        GLA_HIP_TRY(hipStreamSynchronize(s));
        if (herr) {
            set_error("triangular solve: a dependency wait timed out");
            return GLA_ERR_HIP;
        }
        for (int64_t k = 0; k < nrhs; ++k) {
            rr[k] = hn[32 * (k / 16) + k % 16];
            xx[k] = hn[32 * (k / 16) + 16 + k % 16];
        }
        if (it == 0) {
            bb = rr;   // x = 0: r = b
        } else {
            eta = backward_error(rr.data(), xx.data(), bb.data(), nrhs, anorm);
            if (eta <= tol) {
                *outcome = kConverged;
                break;
            }
            if (!std::isfinite(eta) || (it >= 3 && eta > 0.5 * prev2)) {
                *outcome = kStalled;
                break;
            }
            if (corr >= max_iter) break;
            ++corr;
            prev2 = prev1;
            prev1 = eta;
        }
        // x += (L Lᵀ)⁻¹ r: forward solve, then the backward one as a lower
        // solve with Lr on reversed rows, adding into X
        for (int64_t c = 0; c < nch; ++c) {
            const int nr = (int)std::min<int64_t>(16, nrhs - 16 * c);
            GLA_TRY(trsv16<T>(h, s, n, ws.W, n, ws.Dt, ws.Rt + c * ws.npad * 16, ws.Yt, 0, nullptr,
                              0, nr, ws.flags, ++epoch, ws.err));
            GLA_TRY(trsv16<T>(h, s, n, ws.Lr, n, ws.Dtr, ws.Yt, ws.Zt, 1, X + 16 * c * ldx, ldx, nr,
                              ws.flags, ++epoch, ws.err));
        }
    }
    *iters = corr;
    *berr = eta;
    return GLA_OK;
}

}  // namespace

static size_t mixed_ws_bytes(char prec, int64_t n, int64_t nrhs)
{
    return prec == 'd' ? MixedWS<double>::bytes(n, nrhs) : MixedWS<float>::bytes(n, nrhs);
}

// One attempt at precision `prec`: factor, then refine.  *info > 0 when the
// factorisation meets a negative pivot.
template <typename T>
int mixed_attempt(Handle *h, hipStream_t s, char prec, int64_t n, const double *A, int64_t lda,
                  int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx, double tol,
                  int max_iter, double anorm, int *iters, double *berr, int64_t *info, int *outcome)
{
    MixedWS<T> ws;
    ws.carve(h->ws, n, nrhs);
    hipLaunchKernelGGL(fill_int, dim3((unsigned)cdiv(ws.nblk + 1, 256)), dim3(256), 0, s,
                       ws.nblk + 1, ws.flags, 0);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, h->dinfo, 0);
    GLA_LAUNCH_CHECK();
    if constexpr (sizeof(T) == 8) {
        hipLaunchKernelGGL(copy2d_d, g2(n, n), dim3(256), 0, s, n, n, A, lda, ws.W, n);
        GLA_LAUNCH_CHECK();
        GLA_TRY(chol_device<double>(h, s, n, ws.W, n, 2, 1, ws.Linv));
    } else {
        hipLaunchKernelGGL(cvt_lower_d2s, g2(n, n), dim3(256), 0, s, n, A, lda, ws.W, n);
        GLA_LAUNCH_CHECK();
        if (prec == 's')
            GLA_TRY(chol_device<float>(h, s, n, ws.W, n, 2, 1, ws.Linv));
        else if (prec == 'b')
            GLA_TRY(chol_mixed_la<__bf16>(h, s, n, ws.W, n, ws.D64, ws.Linv64, ws.Linv32, ws.Linv,
                                           h->mixed_block));
        else
            GLA_TRY(chol_mixed_la<_Float16>(h, s, n, ws.W, n, ws.D64, ws.Linv64, ws.Linv32, ws.Linv,
                                             h->mixed_block));
    }
    int hinfo = 0;
    GLA_HIP_TRY(hipMemcpyAsync(&hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    GLA_HIP_TRY(hipStreamSynchronize(s));
    *info = hinfo;
    if (hinfo != 0) {
        *iters = 0;
        *berr = INFINITY;
        *outcome = kStalled;
        return GLA_OK;
    }
    GLA_TRY(prepare_solves<T>(s, n, ws));
    int epoch = 0;
    return refine<T>(h, s, n, A, lda, nrhs, B, ldb, X, ldx, ws, anorm, tol, max_iter, epoch, iters,
                     berr, outcome);
}

int mixed_solve(Handle *h, hipStream_t s, char prec, int64_t n, const double *A, int64_t lda,
                int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx, double tol,
                int max_iter, int *iters, double *berr, int64_t *info, char *used)
{
    // precision ladder: 'a' tries BF16, then FP32, then FP64
    const char ladder_a[] = {'b', 's', 'd'};
    const char *ladder = prec == 'a' ? ladder_a : &prec;
    const int nl = prec == 'a' ? 3 : 1;
    size_t need = 0;
    for (int i = 0; i < nl; ++i) need = std::max(need, mixed_ws_bytes(ladder[i], n, nrhs));
    // ‖A‖_F from the lower triangle: column sums in scratch past the layouts
    GLA_TRY(ensure_workspace(h, need + 8 * (size_t)(n + 64)));
    double *cs = reinterpret_cast<double *>(reinterpret_cast<char *>(h->ws) + need);
    std::vector<double> hcs(n);
    hipLaunchKernelGGL(sym_sumsq_cols, dim3((unsigned)n), dim3(256), 0, s, n, A, lda, cs);
    GLA_LAUNCH_CHECK();
    GLA_HIP_TRY(hipMemcpyAsync(hcs.data(), cs, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    GLA_HIP_TRY(hipStreamSynchronize(s));
    double a2 = 0.0;
    for (int64_t j = 0; j < n; ++j) a2 += hcs[j];
    const double anorm = std::sqrt(a2);
    for (int i = 0; i < nl; ++i) {
        const char p = ladder[i];
        int outcome = kMaxIter;
        if (p == 'd')
            GLA_TRY(mixed_attempt<double>(h, s, p, n, A, lda, nrhs, B, ldb, X, ldx, tol, max_iter,
                                          anorm, iters, berr, info, &outcome));
        else
            GLA_TRY(mixed_attempt<float>(h, s, p, n, A, lda, nrhs, B, ldb, X, ldx, tol, max_iter,
                                         anorm, iters, berr, info, &outcome));
        if (used) *used = p;
        if (outcome == kConverged || i == nl - 1) break;
    }
    return GLA_OK;
}

int lp_gemm_entry(char tin, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                  int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc, int mode,
                  hipStream_t s)
{
    if (tin == 'b') return lp_gemm_nt<__bf16>(s, mode, m, n, k, alpha, A, lda, B, ldb, C, ldc);
    return lp_gemm_nt<_Float16>(s, mode, m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

}  // namespace gla

extern "C" {

int gla_chol_solve_mixed_d(gla_handle_t hh, char prec, int64_t n, const double *A, int64_t lda,
                           int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx,
                           double tol, int32_t max_iter, int32_t *iters, double *berr,
                           int64_t *info, char *prec_used, void *stream)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (prec != 'd' && prec != 's' && prec != 'b' && prec != 'h' && prec != 'a') {
        gla::set_error("prec must be one of 'd', 's', 'b', 'h', 'a'");
        return -2;
    }
    if (n < 0) return -3;
    if (lda < std::max<int64_t>(1, n)) return -5;
    if (nrhs < 0) return -6;
    if (ldb < std::max<int64_t>(1, n)) return -8;
    if (ldx < std::max<int64_t>(1, n)) return -10;
    if (max_iter < 0) return -12;
    int it = 0;
    double be = 0.0;
    int64_t inf = 0;
    char used = prec == 'a' ? 'b' : prec;
    int rc = GLA_OK;
    if (n > 0 && nrhs > 0) {
        GLA_HIP_TRY(hipSetDevice(h->device));
        rc = gla::mixed_solve(h, reinterpret_cast<hipStream_t>(stream), prec, n, A, lda, nrhs, B,
                              ldb, X, ldx, tol, max_iter, &it, &be, &inf, &used);
    }
    if (iters) *iters = it;
    if (berr) *berr = be;
    if (info) *info = inf;
    if (prec_used) *prec_used = used;
    return rc;
}

int gla_gemm_lowp_nt(gla_handle_t hh, char tin, int64_t m, int64_t n, int64_t k, float alpha,
                     const float *A, int64_t lda, const float *B, int64_t ldb, float *C,
                     int64_t ldc, void *stream)
{
    auto *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) return -1;
    if (tin != 'b' && tin != 'h') return -2;
    if (alpha != 1.f && alpha != -1.f) { gla::set_error("alpha must be +1 or -1"); return -6; }
    GLA_HIP_TRY(hipSetDevice(h->device));
    return gla::lp_gemm_entry(tin, m, n, k, alpha, A, lda, B, ldb, C, ldc, 0,
                              reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
