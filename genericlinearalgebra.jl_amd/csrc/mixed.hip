// Mixed-precision Cholesky solve with FP64 iterative refinement.
// No counterpart in the reference (its Cholesky runs in one precision,
// src/cholesky.jl).  A x = B for SPD A (n×n, FP64) and nrhs right-hand sides:
//
//   factor   L Lᵀ ≈ A in a working copy W:
//     'd'  FP64 throughout (the plain FP64 path, cholRecursive! on device);
//     's'  FP32 throughout;
//     'b'/'h'  FP32 storage, every 128×128 diagonal block factored in FP64
//          (read and written back in FP32), the panel solve in FP32, and
//          the trailing SYRK/GEMM updates (rankUpdate!, src/juliaBLAS.jl:89)
//          on BF16 / FP16 MFMA with FP32 accumulation
//          (v_mfma_f32_16x16x32_{bf16,f16}) from panels rounded once;
//   refine   r = B − A x in FP64 (symmetric product from A's lower
//            triangle), x += (L Lᵀ)⁻¹ r with the low-precision factor,
//            until the normwise backward error
//            η = max_k ‖r_k‖ / (‖A‖_F ‖x_k‖ + ‖b_k‖) ≤ tol.
//
// The triangular solves are one-launch dataflow solves (trsv16): L y = r
// directly, and Lᵀ z = y as a lower solve with the reversed matrix
// L' = P Lᵀ P (P the reversal), z = P (L'⁻¹ (P y)).
#include <algorithm>
#include <cmath>
#include <type_traits>
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

// Low-precision product on packed operands.  A panel is rounded to TIN once
// and stored row-major with K padded to kPkBK (lp_pack_split): row r holds its K
// values contiguously, which is the MFMA operand layout (a lane reads eight
// consecutive k of one row with one 16-byte load).  The trailing updates of
// the factorisation read each panel row many times, so the rounding and the
// FP32→TIN halving of the bytes are paid once per panel.
constexpr int kPkT = 256, kPkBM = 128, kPkBK = 64, kPkLD = kPkBK + 8;

inline int64_t lp_kpad(int64_t K) { return cdiv(std::max<int64_t>(K, 1), kPkBK) * kPkBK; }

// Multi-term operands.  An FP32 panel X is split into terms of the input
// type, X = X₀ + X₁ (+ X₂) with each term the rounding of what the previous
// ones leave (FP16: two terms, 22 of FP32's 24 significand bits; BF16: three
// terms, all 24), and the update X Xᵀ is formed from the products whose
// size is above the target accuracy as ONE product with a longer K:
//   two terms:   A' = [X₀ X₀ X₁],        B' = [X₀ X₁ X₀]         (3 segments)
//   three terms: A' = [X₀ X₀ X₁ X₀ X₂ X₁], B' = [X₀ X₁ X₀ X₂ X₀ X₁] (6 segments)
// so A'B'ᵀ = X₀X₀ᵀ + X₀X₁ᵀ + X₁X₀ᵀ (+ X₀X₂ᵀ + X₂X₀ᵀ + X₁X₁ᵀ), leaving out
// terms of relative size 2⁻²² (FP16) or 2⁻²⁴ (BF16).  FP16 panels are first
// scaled by a power of two s that brings max|X| to [2¹³, 2¹⁴), so that the
// low terms stay normal; the product is then scaled back by 1/s² (exact).
__host__ __device__ constexpr int lp_nseg(int nterm) { return nterm == 1 ? 1 : (nterm == 2 ? 3 : 6); }

// Scale of an FP16 panel from max|X| (its bits in *maxbits): 2^(14 − e)
// with max|X| < 2^e; 1 without a maximum or for a zero/non-finite one.
__device__ __forceinline__ float lp_scale(const unsigned *maxbits)
{
    if (!maxbits) return 1.f;
    const float m = __uint_as_float(*maxbits);
    if (!(m > 0.f) || !(m < INFINITY)) return 1.f;
    int e;
    (void)frexpf(m, &e);
    return ldexpf(1.f, 14 - e);
}

__global__ void lp_maxabs_kernel(int64_t rows, int64_t K, const float *src, int64_t lds,
                                 unsigned *maxbits)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.f;
    if (r < rows)
        for (int64_t k = blockIdx.y; k < K; k += gridDim.y) m = fmaxf(m, fabsf(src[r + k * lds]));
    // a wave's maximum, then one atomic per wave (non-negative floats order
    // as their bit patterns)
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(maxbits, __float_as_uint(m));
}

__global__ void lp_zero_u32(unsigned *p) { *p = 0u; }

// A' and B' (row-major, pitch nseg·Kp) of the FP32 panel src (rows × K),
// NT terms of TIN, scaled by lp_scale(maxbits).  Null dA or dB: that
// operand's pattern is not written (a general product packs its two factors
// separately).
template <typename TIN, int NT>
__global__ void lp_pack_split_kernel(int64_t rows, int64_t K, int64_t Kp, const float *src,
                                     int64_t lds, TIN *dA, TIN *dB, const unsigned *maxbits)
{
    typedef typename LpMfma<TIN>::v8 v8;
    constexpr int NS = lp_nseg(NT);
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float sc = lp_scale(maxbits);
    const int64_t P = NS * Kp;
    for (int64_t k0 = 8 * (int64_t)blockIdx.y; k0 < Kp; k0 += 8 * (int64_t)gridDim.y) {
        v8 t[NT];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float x = k0 + e < K ? src[r + (k0 + e) * lds] * sc : 0.f;
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                const TIN v = (TIN)x;
                t[q][e] = v;
                x -= (float)v;
            }
        }
        // segment order of A' / B' (see above)
        if (dA) {
            TIN *a = dA + r * P + k0;
            *reinterpret_cast<v8 *>(a) = t[0];
            if constexpr (NT >= 2) {
                *reinterpret_cast<v8 *>(a + Kp) = t[0];
                *reinterpret_cast<v8 *>(a + 2 * Kp) = t[1];
            }
            if constexpr (NT >= 3) {
                *reinterpret_cast<v8 *>(a + 3 * Kp) = t[0];
                *reinterpret_cast<v8 *>(a + 4 * Kp) = t[2];
                *reinterpret_cast<v8 *>(a + 5 * Kp) = t[1];
            }
        }
        if (dB && dB != dA) {
            TIN *b = dB + r * P + k0;
            *reinterpret_cast<v8 *>(b) = t[0];
            if constexpr (NT >= 2) {
                *reinterpret_cast<v8 *>(b + Kp) = t[1];
                *reinterpret_cast<v8 *>(b + 2 * Kp) = t[0];
            }
            if constexpr (NT >= 3) {
                *reinterpret_cast<v8 *>(b + 3 * Kp) = t[2];
                *reinterpret_cast<v8 *>(b + 4 * Kp) = t[0];
                *reinterpret_cast<v8 *>(b + 5 * Kp) = t[1];
            }
        }
    }
}

// One packed panel of the factorisation: A' and B' (the same buffer for a
// single term), pitch nseg·lp_kpad(K), and the FP16 scale word.
template <typename TIN>
struct LpPanel {
    TIN *a = nullptr, *b = nullptr;
    unsigned *maxbits = nullptr;   // null: no scaling (BF16)
    int nterm = 1;
    int64_t pitch(int64_t K) const { return lp_nseg(nterm) * lp_kpad(K); }
};

// max|src| over rows × K into *maxbits (atomic maximum: callers zero it
// first, and may accumulate several operands into one word).
inline int lp_maxabs(hipStream_t s, int64_t rows, int64_t K, const float *src, int64_t lds,
                     unsigned *maxbits)
{
    if (rows <= 0 || K <= 0) return GLA_OK;
    hipLaunchKernelGGL(lp_maxabs_kernel, dim3((unsigned)cdiv(rows, 256),
                                              (unsigned)std::min<int64_t>(K, 64)),
                       dim3(256), 0, s, rows, K, src, lds, maxbits);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// Packs src into dA (A' pattern) and/or dB (B' pattern) with nterm terms.
template <typename TIN>
int lp_pack_split(hipStream_t s, int nterm, int64_t rows, int64_t K, const float *src,
                  int64_t lds, TIN *dA, TIN *dB, const unsigned *maxbits)
{
    if (rows <= 0) return GLA_OK;
    const int64_t Kp = lp_kpad(K);
    const dim3 grid((unsigned)cdiv(rows, 256), (unsigned)std::min<int64_t>(Kp / 8, 65535));
    if (nterm == 1)
        hipLaunchKernelGGL((lp_pack_split_kernel<TIN, 1>), grid, dim3(256), 0, s, rows, K, Kp, src,
                           lds, dA, dB, maxbits);
    else if (nterm == 2)
        hipLaunchKernelGGL((lp_pack_split_kernel<TIN, 2>), grid, dim3(256), 0, s, rows, K, Kp, src,
                           lds, dA, dB, maxbits);
    else
        hipLaunchKernelGGL((lp_pack_split_kernel<TIN, 3>), grid, dim3(256), 0, s, rows, K, Kp, src,
                           lds, dA, dB, maxbits);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// The packed A' and B' of one panel of the factorisation (its SYRK update).
template <typename TIN>
int lp_pack_panel(hipStream_t s, int64_t rows, int64_t K, const float *src, int64_t lds,
                  const LpPanel<TIN> &pp)
{
    if (rows <= 0) return GLA_OK;
    if (pp.maxbits) {
        hipLaunchKernelGGL(lp_zero_u32, dim3(1), dim3(1), 0, s, pp.maxbits);
        GLA_LAUNCH_CHECK();
        GLA_TRY(lp_maxabs(s, rows, K, src, lds, pp.maxbits));
    }
    return lp_pack_split<TIN>(s, pp.nterm, rows, K, src, lds, pp.a, pp.b, pp.maxbits);
}

struct LpPk {
    int64_t M, N, Kp;
    const void *A;   // M×Kp, TIN, row-major
    const void *B;   // N×Kp, TIN, row-major
    float *C;
    int64_t ldc;
    float alpha;     // C <- C + alpha A Bᵀ (divided by lp_scale(maxbits)²)
    const unsigned *maxbits;
    int mode;        // 0 all, 1 lower (row >= col)
    int tiles_m, tiles_n, total, per_xcd;
};

// Tile order.  Tiles are grouped in kPkST×kPkST super-tiles (1024×1024 of
// C), super-tiles numbered column by column (in mode 1 only those on or
// below the diagonal), tiles column by column inside each.  Consecutive
// tile numbers go to one XCD, whose ~64 resident workgroups then share one
// super-tile's 1024 rows of A and of B in its L2.  In mode 1 the tiles of a
// diagonal super-tile that lie wholly above the diagonal are numbered too
// and exit at once.
constexpr int kPkST = 8;

__host__ __device__ inline int lp_st_first(int J, int mode) { return mode ? J : 0; }

// number of tiles in super-tile column J
__host__ __device__ inline int64_t lp_st_col_tiles(int J, int tiles_m, int tiles_n, int mode)
{
    const int wd = min(kPkST, tiles_n - J * kPkST);
    const int rows = tiles_m - lp_st_first(J, mode) * kPkST;
    return rows > 0 ? (int64_t)wd * rows : 0;
}

__device__ __forceinline__ void lp_tile_of(int64_t tid, int tiles_m, int tiles_n, int mode, int &tm,
                                           int &tn)
{
    const int SN = (tiles_n + kPkST - 1) / kPkST;
    int J = 0;
    for (; J < SN; ++J) {
        const int64_t c = lp_st_col_tiles(J, tiles_m, tiles_n, mode);
        if (tid < c) break;
        tid -= c;
    }
    const int wd = min(kPkST, tiles_n - J * kPkST);
    const int I = lp_st_first(J, mode) + (int)(tid / ((int64_t)kPkST * wd));
    const int inner = (int)(tid - (int64_t)(I - lp_st_first(J, mode)) * kPkST * wd);
    const int h = min(kPkST, tiles_m - I * kPkST);
    tn = J * kPkST + inner / h;
    tm = I * kPkST + inner % h;
}

// C (M×N, FP32, column-major) += alpha · A Bᵀ on packed TIN operands, FP32
// accumulation.  128×128 tiles, four waves of 64×64, K in steps of 64
// staged through double-buffered LDS with the next step's loads in flight
// during the MFMAs.  Tiles are numbered column by column (lower tiles only
// in mode 1) and dealt to the XCDs in contiguous runs, so neighbouring
// tiles share an L2.  The MFMA is issued as Cᵀ = B Aᵀ: a lane's results
// then run down a column of C (16 consecutive rows per 16 lanes).
template <typename TIN>
__global__ __launch_bounds__(kPkT, 2) void lp_gemm_pk_kernel(LpPk p)
{
    typedef typename LpMfma<TIN>::v8 v8;
    __shared__ __attribute__((aligned(16))) TIN As[2][kPkBM * kPkLD];
    __shared__ __attribute__((aligned(16))) TIN Bs[2][kPkBM * kPkLD];

    const int b = blockIdx.x;
    const int tid = (b % 8) * p.per_xcd + b / 8;
    if (tid >= p.total) return;
    int tm, tn;
    lp_tile_of(tid, p.tiles_m, p.tiles_n, p.mode, tm, tn);
    if (p.mode == 1 && tm < tn) return;          // wholly above the diagonal
    const int64_t m0 = (int64_t)tm * kPkBM, n0 = (int64_t)tn * kPkBM;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const TIN *__restrict__ A = reinterpret_cast<const TIN *>(p.A);
    const TIN *__restrict__ B = reinterpret_cast<const TIN *>(p.B);
    const int64_t Kp = p.Kp;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    // staging: 16-byte chunk c = t + 256u (u < 4) of the 128×64 slab is row
    // c / 8, k 8·(c % 8): eight threads read one row's 128 contiguous bytes
    v8 ra[4], rb[4];
    auto gload = [&](const TIN *X, int64_t x0, int64_t xlim, int64_t k0, v8 *r) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = t + kPkT * u, row = c >> 3, kc = c & 7;
            const int64_t gx = x0 + row;
            if (gx < xlim)
                r[u] = *reinterpret_cast<const v8 *>(X + gx * Kp + k0 + 8 * kc);
            else
#pragma unroll
                for (int e = 0; e < 8; ++e) r[u][e] = (TIN)0.f;
        }
    };
    auto lstore = [&](TIN *S, const v8 *r) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = t + kPkT * u, row = c >> 3, kc = c & 7;
            *reinterpret_cast<v8 *>(&S[row * kPkLD + 8 * kc]) = r[u];
        }
    };

    const int nk = (int)(Kp / kPkBK);
    gload(A, m0, p.M, 0, ra);
    gload(B, n0, p.N, 0, rb);
    lstore(As[0], ra);
    lstore(Bs[0], rb);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            gload(A, m0, p.M, (int64_t)(kt + 1) * kPkBK, ra);
            gload(B, n0, p.N, (int64_t)(kt + 1) * kPkBK, rb);
        }
#pragma unroll
        for (int kk = 0; kk < kPkBK / 32; ++kk) {
            v8 a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *reinterpret_cast<const v8 *>(
                    &As[cur][(wm * 64 + i * 16 + (lane & 15)) * kPkLD + 32 * kk + 8 * (lane >> 4)]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bb[j] = *reinterpret_cast<const v8 *>(
                    &Bs[cur][(wn * 64 + j * 16 + (lane & 15)) * kPkLD + 32 * kk + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = LpMfma<TIN>::mma(bb[j], a[i], acc[i][j]);
        }
        if (more) {
            lstore(As[cur ^ 1], ra);
            lstore(Bs[cur ^ 1], rb);
        }
        __syncthreads();
    }

    // acc[i][j][r] = C(m0 + 64wm + 16i + lane%16, n0 + 64wn + 16j + 4(lane/16) + r)
    float alpha = p.alpha;
    if (p.maxbits) {
        const float sc = lp_scale(p.maxbits);
        alpha = alpha / (sc * sc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t gr = m0 + wm * 64 + i * 16 + (lane & 15);
        if (gr >= p.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gc = n0 + wn * 64 + j * 16 + 4 * (lane >> 4) + r;
                if (gc < p.N && (p.mode == 0 || gr >= gc)) {
                    float *c = p.C + gr + gc * p.ldc;
                    *c = *c + alpha * acc[i][j][r];
                }
            }
    }
}

// C (M×N) += alpha · A Bᵀ with A, B packed by lp_pack_split (pitch lp_kpad(K)).
template <typename TIN>
int lp_gemm_pk(hipStream_t s, int mode, int64_t M, int64_t N, int64_t K, float alpha,
               const TIN *A, const TIN *B, float *C, int64_t ldc,
               const unsigned *maxbits = nullptr)
{
    if (M <= 0 || N <= 0) return GLA_OK;
    LpPk p{};
    p.M = M;
    p.N = N;
    p.Kp = lp_kpad(K);
    p.maxbits = maxbits;
    p.A = A;
    p.B = B;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.mode = mode;
    p.tiles_m = (int)cdiv(M, kPkBM);
    p.tiles_n = (int)cdiv(N, kPkBM);
    int64_t tot = 0;
    for (int J = 0; J < (int)cdiv(p.tiles_n, kPkST); ++J)
        tot += lp_st_col_tiles(J, p.tiles_m, p.tiles_n, mode);
    p.total = (int)tot;
    if (p.total == 0) return GLA_OK;
    p.per_xcd = (int)cdiv(p.total, 8);
    hipLaunchKernelGGL((lp_gemm_pk_kernel<TIN>), dim3(8 * p.per_xcd), dim3(kPkT), 0, s, p);
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
// blocks and low-precision trailing updates (n ≤ the look-ahead block
// width).  pk: room for the packed panel, n × lp_kpad(128) TIN.
template <typename TIN>
int chol_mixed(Handle *h, hipStream_t s, int64_t n, float *W, int64_t ldw, float *Linv32,
               const LpPanel<TIN> &pk)
{
    const int nb = 128;
    const int64_t nsteps = cdiv(n, nb);
    int64_t k = 0;
    for (int64_t step = 0; step < nsteps; ++step, k += nb) {
        const int kb = (int)std::min<int64_t>(nb, n - k);
        const int64_t below = n - k - kb;
        float *Wkk = W + k + k * ldw;
        float *W21 = Wkk + kb;
        // diagonal block factored in FP64 (read and written back in FP32)
        GLA_TRY(chol_factor_block_f32io(s, kb, Wkk, ldw, h->dinfo, k, Linv32));
        GLA_TRY(chol_solve_block<float>(s, kb, Wkk, ldw, Linv32, W21, ldw, below));
        if (below == 0) break;
        const int64_t kb2 = std::min<int64_t>(nb, below);
        const int64_t rest = below - kb2;
        const int64_t P = pk.pitch(kb);
        GLA_TRY(lp_pack_panel<TIN>(s, below, kb, W21, ldw, pk));
        GLA_TRY(lp_gemm_pk<TIN>(s, 1, below, kb2, P, -1.f, pk.a, pk.b, W21 + kb * ldw, ldw,
                                pk.maxbits));
        if (rest > 0)
            GLA_TRY(lp_gemm_pk<TIN>(s, 1, rest, rest, P, -1.f, pk.a + kb2 * P, pk.b + kb2 * P,
                                    W21 + kb2 + (kb + kb2) * ldw, ldw, pk.maxbits));
    }
    return GLA_OK;
}

// Wide-block form (block width NB): the NB×NB diagonal block by the
// 128-wide driver above on the caller's stream, the panel below it by the
// FP32 recursive solve, then the look-ahead trailing updates of
// cholBlocked!'s loop (src/cholesky.jl:17-35) with K = NB on the
// low-precision MFMA GEMM: the next block column on the caller's stream,
// the rest on the side stream.  The panel is packed once per step into
// pk[step % 2] (each n × lp_kpad(NB) TIN): the side stream may still read
// the previous step's buffer while the caller's stream packs the next.
template <typename TIN>
int chol_mixed_la(Handle *h, hipStream_t s, int64_t n, float *W, int64_t ldw, float *Linv32,
                  float *Linv, int NB, const LpPanel<TIN> &pk0, const LpPanel<TIN> &pk1)
{
    if (n <= NB) return chol_mixed<TIN>(h, s, n, W, ldw, Linv32, pk0);
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
        const LpPanel<TIN> &pk = (step & 1) ? pk1 : pk0;
        GLA_TRY(chol_mixed<TIN>(h, s, kb, Wkk, ldw, Linv32, pk));
        if (below == 0) break;
        // every fourth block, stop early once a pivot has failed: a
        // low-precision attempt that breaks down is abandoned for the next
        // precision without factoring the rest
        if ((step & 3) == 3) {
            int hinfo = 0;
            GLA_HIP_TRY(hipMemcpyAsync(&hinfo, h->dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
            GLA_HIP_TRY(hipStreamSynchronize(s));
            if (hinfo != 0) break;
        }
        GLA_TRY(tile_inverses<float>(s, kb, Wkk, ldw, Linv));
        GLA_TRY(trsm_rlt<float>(s, below, kb, Wkk, ldw, Linv, W21, ldw));
        const int64_t kb2 = std::min<int64_t>(NB, below);
        const int64_t rest = below - kb2;
        const int64_t P = pk.pitch(kb);
        GLA_TRY(lp_pack_panel<TIN>(s, below, kb, W21, ldw, pk));
        // as in the FP64 look-ahead driver, the side stream's trailing update
        // starts after the next block column's (chol_syrk_late) or with it
        const bool late = h->chol_syrk_late != 0;
        if (overlap && rest > 0 && !late) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (overlap && step > 0) GLA_HIP_TRY(hipStreamWaitEvent(s, h->ev_step[1], 0));
        GLA_TRY(lp_gemm_pk<TIN>(s, 1, below, kb2, P, -1.f, pk.a, pk.b, W21 + kb * ldw, ldw,
                                pk.maxbits));
        if (overlap && rest > 0 && late) GLA_HIP_TRY(hipEventRecord(h->ev_step[0], s));
        if (rest > 0) {
            hipStream_t us = overlap ? h->side : s;
            if (overlap) GLA_HIP_TRY(hipStreamWaitEvent(h->side, h->ev_step[0], 0));
            GLA_TRY(lp_gemm_pk<TIN>(us, 1, rest, rest, P, -1.f, pk.a + kb2 * P, pk.b + kb2 * P,
                                    W21 + kb2 + (kb + kb2) * ldw, ldw, pk.maxbits));
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
    T *W, *Lr, *Linv, *Linvr, *Dt, *Dtr, *Mt, *Mtr, *Rt, *Yt, *Zt;
    double *Xt, *nrm;
    float *Linv32;
    int *flags, *err;
    // packed low-precision panels (A' and B' of two alternating steps),
    // pk_bytes each, and their FP16 scale words
    void *pk0a, *pk0b, *pk1a, *pk1b;
    unsigned *maxbits;
    int64_t nblk, npad, nchunk;

    // pk_bytes: one packed panel of the BF16/FP16 factorisation (0 otherwise)
    static size_t bytes(int64_t n, int64_t nrhs, size_t pk_bytes)
    {
        const int64_t nblk = cdiv(n, 128), npad = nblk * 128, nch = cdiv(nrhs, 16);
        const size_t es = sizeof(T), r = 255;
        auto al = [&](size_t b) { return (b + r) / 256 * 256; };
        return 2 * al(es * (size_t)n * n) + 2 * al(es * tile_inverse_elems(n)) +
               4 * al(es * (size_t)npad * 128) + al(es * (size_t)npad * 16 * nch) +
               2 * al(es * (size_t)npad * 16) + al(8 * symv_scratch(n)) +
               al(8 * (size_t)(32 * nch + n + 64)) +
               al(4 * 8 * 256) + al(4 * (size_t)(nblk + 64)) + 4 * al(pk_bytes) + 4096;
    }
    void carve(void *base, int64_t n, int64_t nrhs, size_t pk_bytes)
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
        Mt = reinterpret_cast<T *>(take(es * (size_t)npad * 128));
        Mtr = reinterpret_cast<T *>(take(es * (size_t)npad * 128));
        Rt = reinterpret_cast<T *>(take(es * (size_t)npad * 16 * nchunk));
        Yt = reinterpret_cast<T *>(take(es * (size_t)npad * 16));
        Zt = reinterpret_cast<T *>(take(es * (size_t)npad * 16));
        Xt = reinterpret_cast<double *>(take(8 * symv_scratch(n)));
        nrm = reinterpret_cast<double *>(take(8 * (size_t)(32 * nchunk + n + 64)));
        Linv32 = reinterpret_cast<float *>(take(4 * 8 * 256));
        flags = reinterpret_cast<int *>(take(4 * (size_t)(nblk + 64)));
        err = flags + nblk;
        maxbits = reinterpret_cast<unsigned *>(flags + nblk + 8);
        pk0a = pk_bytes ? take(pk_bytes) : nullptr;
        pk0b = pk_bytes ? take(pk_bytes) : nullptr;
        pk1a = pk_bytes ? take(pk_bytes) : nullptr;
        pk1b = pk_bytes ? take(pk_bytes) : nullptr;
    }
};

// Precisions with low-precision trailing updates: 'b' / 'h' one BF16 /
// FP16 term, 'B' BF16 in three terms, 'H' FP16 in two (scaled) terms.
inline bool is_lowp(char prec) { return prec == 'b' || prec == 'h' || prec == 'B' || prec == 'H'; }
inline int lowp_terms(char prec) { return prec == 'B' ? 3 : (prec == 'H' ? 2 : 1); }

// Bytes of one packed operand (A' or B') for precision prec at block width NB.
inline size_t pk_bytes_of(char prec, int64_t n, int NB)
{
    return is_lowp(prec) ? 2 * (size_t)n * lp_nseg(lowp_terms(prec)) * (size_t)lp_kpad(NB) : 0;
}

template <typename TIN>
int chol_mixed_run(Handle *h, hipStream_t s, char prec, int64_t n, float *W, float *Linv32,
                   float *Linv, MixedWS<float> &ws)
{
    LpPanel<TIN> p0, p1;
    p0.nterm = p1.nterm = lowp_terms(prec);
    p0.a = reinterpret_cast<TIN *>(ws.pk0a);
    p1.a = reinterpret_cast<TIN *>(ws.pk1a);
    p0.b = p0.nterm > 1 ? reinterpret_cast<TIN *>(ws.pk0b) : p0.a;
    p1.b = p1.nterm > 1 ? reinterpret_cast<TIN *>(ws.pk1b) : p1.a;
    if constexpr (sizeof(TIN) == 2 && !std::is_same<TIN, __bf16>::value) {
        p0.maxbits = ws.maxbits;
        p1.maxbits = ws.maxbits + 1;
    }
    return chol_mixed_la<TIN>(h, s, n, W, n, Linv32, Linv, h->mixed_block, p0, p1);
}

// Solve structures of the factor in ws.W: tile inverses, diagonal-block
// inverses, and the same for the reversed Lr = J Lᵀ J (backward solve).
template <typename T>
int prepare_solves(hipStream_t s, int64_t n, MixedWS<T> &ws)
{
    GLA_TRY(tile_inverses<T>(s, n, ws.W, n, ws.Linv));
    GLA_TRY(diag_block_inverses<T>(s, n, ws.W, n, ws.Linv, ws.Dt));
    GLA_TRY(diag_sub<T>(s, n, ws.W, n, ws.Dt, ws.Mt));
    hipLaunchKernelGGL((reverse_transpose_lower<T>), g2(n, n), dim3(256), 0, s, n, ws.W, n, ws.Lr,
                       n);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses<T>(s, n, ws.Lr, n, ws.Linvr));
    GLA_TRY(diag_block_inverses<T>(s, n, ws.Lr, n, ws.Linvr, ws.Dtr));
    return diag_sub<T>(s, n, ws.Lr, n, ws.Dtr, ws.Mtr);
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
            GLA_TRY(trsv16<T>(h, s, n, ws.W, n, ws.Dt, ws.Mt, ws.Rt + c * ws.npad * 16, ws.Yt, 0,
                              nullptr, 0, nr, ws.flags, ++epoch, ws.err));
            GLA_TRY(trsv16<T>(h, s, n, ws.Lr, n, ws.Dtr, ws.Mtr, ws.Yt, ws.Zt, 1, X + 16 * c * ldx,
                              ldx, nr, ws.flags, ++epoch, ws.err));
        }
    }
    *iters = corr;
    *berr = eta;
    return GLA_OK;
}

}  // namespace

static size_t mixed_ws_bytes(char prec, int64_t n, int64_t nrhs, int NB)
{
    return prec == 'd' ? MixedWS<double>::bytes(n, nrhs, 0)
                       : MixedWS<float>::bytes(n, nrhs, pk_bytes_of(prec, n, NB));
}

// One attempt at precision `prec`: factor, then refine.  *info > 0 when the
// factorisation meets a negative pivot.
template <typename T>
int mixed_attempt(Handle *h, hipStream_t s, char prec, int64_t n, const double *A, int64_t lda,
                  int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx, double tol,
                  int max_iter, double anorm, int *iters, double *berr, int64_t *info, int *outcome)
{
    MixedWS<T> ws;
    ws.carve(h->ws, n, nrhs, pk_bytes_of(prec, n, h->mixed_block));
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
        else if (prec == 'b' || prec == 'B')
            GLA_TRY(chol_mixed_run<__bf16>(h, s, prec, n, ws.W, ws.Linv32, ws.Linv, ws));
        else
            GLA_TRY(chol_mixed_run<_Float16>(h, s, prec, n, ws.W, ws.Linv32, ws.Linv, ws));
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
    // precision ladder: 'a' tries FP32, then FP64.  On gfx950 at n = 16384
    // the BF16/FP16 factorisations are no faster than FP32 (the FP64
    // diagonal-block chain, not the trailing MFMA updates, sets their pace)
    // and need two to four times the refinement sweeps, so they are
    // selected explicitly ('b', 'h') rather than tried first.
    const char ladder_a[] = {'s', 'd'};
    const char *ladder = prec == 'a' ? ladder_a : &prec;
    const int nl = prec == 'a' ? 2 : 1;
    size_t need = 0;
    for (int i = 0; i < nl; ++i)
        need = std::max(need, mixed_ws_bytes(ladder[i], n, nrhs, h->mixed_block));
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

template <typename TIN>
int lp_gemm_packed_entry(Handle *h, int nterm, bool scaled, int64_t m, int64_t n, int64_t k,
                         float alpha, const float *A, int64_t lda, const float *B, int64_t ldb,
                         float *C, int64_t ldc, hipStream_t s)
{
    if (m <= 0 || n <= 0) return GLA_OK;
    const int64_t P = lp_nseg(nterm) * lp_kpad(k);
    const size_t pa = ((size_t)m * P * sizeof(TIN) + 255) / 256 * 256;
    const size_t pb = ((size_t)n * P * sizeof(TIN) + 255) / 256 * 256;
    GLA_TRY(ensure_workspace(h, pa + pb + 256));
    TIN *Ap = reinterpret_cast<TIN *>(h->ws);
    TIN *Bp = reinterpret_cast<TIN *>(reinterpret_cast<char *>(h->ws) + pa);
    unsigned *mb = scaled ? reinterpret_cast<unsigned *>(reinterpret_cast<char *>(h->ws) + pa + pb)
                          : nullptr;
    if (mb) {
        // one scale for both factors, from the larger of their maxima
        hipLaunchKernelGGL(lp_zero_u32, dim3(1), dim3(1), 0, s, mb);
        GLA_LAUNCH_CHECK();
        GLA_TRY(lp_maxabs(s, m, k, A, lda, mb));
        GLA_TRY(lp_maxabs(s, n, k, B, ldb, mb));
    }
    GLA_TRY(lp_pack_split<TIN>(s, nterm, m, k, A, lda, Ap, nullptr, mb));
    GLA_TRY(lp_pack_split<TIN>(s, nterm, n, k, B, ldb, nullptr, Bp, mb));
    return lp_gemm_pk<TIN>(s, 0, m, n, P, alpha, Ap, Bp, C, ldc, mb);
}

int lp_gemm_entry(Handle *h, char tin, int64_t m, int64_t n, int64_t k, float alpha,
                  const float *A, int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc,
                  hipStream_t s)
{
    if (tin == 'b' || tin == 'B')
        return lp_gemm_packed_entry<__bf16>(h, lowp_terms(tin), false, m, n, k, alpha, A, lda, B,
                                            ldb, C, ldc, s);
    return lp_gemm_packed_entry<_Float16>(h, lowp_terms(tin), tin == 'H', m, n, k, alpha, A, lda,
                                          B, ldb, C, ldc, s);
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
    if (prec != 'd' && prec != 's' && prec != 'b' && prec != 'h' && prec != 'B' &&
        prec != 'H' && prec != 'a') {
        gla::set_error("prec must be one of 'd', 's', 'b', 'h', 'B', 'H', 'a'");
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
    char used = prec == 'a' ? 's' : prec;
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
    if (tin != 'b' && tin != 'h' && tin != 'B' && tin != 'H') return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    if (lda < std::max<int64_t>(1, m)) return -8;
    if (ldb < std::max<int64_t>(1, n)) return -10;
    if (ldc < std::max<int64_t>(1, m)) return -12;
    GLA_HIP_TRY(hipSetDevice(h->device));
    return gla::lp_gemm_entry(h, tin, m, n, k, alpha, A, lda, B, ldb, C, ldc,
                              reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
