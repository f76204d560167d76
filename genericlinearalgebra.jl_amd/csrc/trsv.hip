// Solves with a Cholesky factor for a few right-hand sides, and the
// symmetric residual, as used by the mixed-precision refinement.  No
// counterpart in the reference (its Cholesky stops at the factor,
// src/cholesky.jl:3-55); the solve is what `L \ (L' \ b)` would do there.
//
// trsv16: Y = L⁻¹ R for up to 16 right-hand sides, held row-major (row i at
// [16 i]), in one launch.  The rows are cut into 128-row blocks; block b is
//   Y_b = L_bb⁻¹ (R_b − Σ_{j<b} L_bj Y_j),
// each workgroup owns blocks b ≡ blockIdx (mod grid) and streams the tiles
// L_bj (j = 0, 1, …) through MFMA as soon as Y_j is published: a block's
// result is written with device-coherent stores and announced by a flag
// (epoch value) that dependent workgroups poll, with a bounded spin.  The
// diagonal solve is one more 128×128×16 MFMA product with the explicit
// inverse L_bb⁻¹ (from diag_block_inverses); the last link is folded into
// it, Y_b = L_bb⁻¹(R_b − Σ_{j<b−1} L_bj Y_j) − (L_bb⁻¹ L_{b,b−1}) Y_{b−1}, so
// only one product follows the arrival of Y_{b−1}.  The grid never exceeds one
// workgroup per CU and every wait is on a block of a smaller index, so all
// waiters are resident and the launch cannot deadlock.
//
// symv_residual: R = B − A X for symmetric A given by its lower triangle
// (FP64), 16 right-hand sides, with the column sums of squares of R and X
// for the backward error.
#include <algorithm>

#include "gla_coherent.hpp"
#include "gla_internal.hpp"

namespace gla {

namespace {

constexpr int kSB = 128;                  // rows per block
constexpr int kSR = 16;                   // right-hand sides per pass
constexpr int kST = 256;                  // threads per workgroup (residual)
constexpr int kTT = 256;                  // threads per workgroup (solve: four waves)
constexpr unsigned kSolveSpin = 1u << 25;

template <typename T>
struct SMfma;
template <>
struct SMfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct SMfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// One wave's part of a 128-column block of L, KQ column quads at a time
// (a "unit": rows r0 + 32w + 16t + lc, t < 2, columns 4(q0 + kq) + lk,
// kq < KQ) — the MFMA A-operand layout.  Units are loaded one ahead of
// their use, so that only the wait, the Y block and the MFMAs remain on
// the dependency chain.  Rows at or past n read as zero.
template <typename T, int KQ>
__device__ __forceinline__ void load_unit(const T *__restrict__ L, int64_t ldl, int64_t n, int64_t r0,
                                          int q0, int w, int lc, int lk, T (&u)[2][KQ])
{
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int64_t row = r0 + 32 * w + 16 * t + lc;
        const bool ok = row < n;
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq)
            u[t][kq] = ok ? L[row + (int64_t)(4 * (q0 + kq) + lk) * ldl] : T(0);
    }
}

// Block b of the solve:  Y_b = L_bb⁻¹ (R_b − Σ_{j<b−1} L_bj Y_j) − M_b Y_{b−1}
// with M_b = L_bb⁻¹ L_{b,b−1} precomputed (diag_sub), so that once Y_{b−1}
// is published only one 128×128×16 product stands between it and Y_b.
// Four waves of 32 rows (two 16-row MFMA tiles) each; a block of L is
// streamed in 32/KQ units (KQ = 16 for FP64 keeps two units and the
// diagonal inverse's rows within the register file).
template <typename T>
__global__ __launch_bounds__(kTT, 1) void trsv16_kernel(int64_t n, int64_t nblk,
                                                        const T *__restrict__ L, int64_t ldl,
                                                        const T *__restrict__ Dt,
                                                        const T *__restrict__ Mt, int64_t npad,
                                                        const T *__restrict__ R, T *Y, int rev,
                                                        double *__restrict__ X, int64_t ldx,
                                                        int nrhs, int *flags, int epoch, int *err)
{
    typedef SMfma<T> M;
    typedef typename M::acc_t acc_t;
    constexpr int KQ = sizeof(T) == 8 ? 16 : 32;   // column quads per unit
    constexpr int U = (kSB / 4) / KQ;               // units per block
    __shared__ T Ys[2][kSB * kSR];
    __shared__ T Ss[kSB * kSR];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lc = lane & 15, lk = lane >> 4;
    static_assert(kSB * kSR == 8 * kTT, "eight Y elements per thread");

    // wait for block j, then stage Y_j in LDS buffer ys
    auto fetch_y = [&](int64_t j, T *ys) {
        if (threadIdx.x == 0) {
            unsigned spins = 0;
            while (load_coherent(flags + j) < epoch) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSolveSpin) {
                    store_coherent(err, 1);
                    break;
                }
            }
        }
        __syncthreads();
        const T *src = Y + j * (kSB * kSR);
        const T *p[8];
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] = src + threadIdx.x + q * kTT;
        load8_coherent(p, v);
#pragma unroll
        for (int q = 0; q < 8; ++q) ys[threadIdx.x + q * kTT] = v[q];
        __syncthreads();
    };
    // c += unit · Y[q0·4 .. q0·4 + 4KQ) rows of ys
    auto prod = [&](const T (&u)[2][KQ], int q0, const T *ys, acc_t (&c)[2]) {
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) {
            const T bv = ys[(4 * (q0 + kq) + lk) * kSR + lc];
#pragma unroll
            for (int t = 0; t < 2; ++t) c[t] = M::mma(u[t][kq], bv, c[t]);
        }
    };

    for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int64_t r0 = b * kSB;
        acc_t acc[2];                                 // Σ_{j<b−1} L_bj Y_j
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[t] = acc_t{T(0), T(0), T(0), T(0)};
        const int64_t nl = b > 0 ? b - 1 : 0;   // blocks applied before the diagonal solve
        const int64_t nu = nl * U;              // units applied before it
        // unit e < nu is block e / U, quads KQ·(e % U); unit nu + h is M_b's h-th
        auto load_e = [&](int64_t e, T (&u)[2][KQ]) {
            if (e < nu)
                load_unit<T, KQ>(L + (e / U) * kSB * ldl, ldl, n, r0, KQ * (int)(e % U), w, lc, lk,
                                 u);
            else
                load_unit<T, KQ>(Mt, npad, n, r0, KQ * (int)(e - nu), w, lc, lk, u);
        };
        T ua[2][KQ], ub[2][KQ];
        if (b > 0) load_e(0, ua);
        auto link = [&](int64_t e, T (&use)[2][KQ], T (&nxt)[2][KQ]) {
            T *ys = Ys[(e / U) & 1];
            if (e % U == 0) fetch_y(e / U, ys);
            load_e(e + 1, nxt);
            prod(use, KQ * (int)(e % U), ys, acc);
        };
        int64_t e = 0;
        for (; e + 2 <= nu; e += 2) {
            link(e, ua, ub);
            link(e + 1, ub, ua);
        }
        if (e < nu) link(e++, ua, ub);
        // now M_b's first unit is in ua (nu even) or ub (nu odd); with two
        // units, the second goes into the other buffer
        const bool m_in_b = (nu & 1) != 0;
        if (U == 2 && b > 0) {
            if (m_in_b) load_e(nu + 1, ua);
            else load_e(nu + 1, ub);
        }
        // y = L_bb⁻¹ S, S = R_b − acc; this wave's rows of L_bb⁻¹
        // (L_bb⁻¹[r][k] = Dt_b[k + 128 r]) load while S is staged
        T dv[2][kSB / 4];
        {
            const T *D = Dt + b * (kSB * kSB);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int rr = 32 * w + 16 * t + lc;
#pragma unroll
                for (int kq = 0; kq < kSB / 4; ++kq) dv[t][kq] = D[4 * kq + lk + (int64_t)rr * kSB];
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + 32 * w + 16 * t + M::row(lane, r);
                const T rv = row < n ? R[(rev ? n - 1 - row : row) * kSR + lc] : T(0);
                Ss[(32 * w + 16 * t + M::row(lane, r)) * kSR + lc] = rv - acc[t][r];
            }
        __syncthreads();
        acc_t y[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) y[t] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int kq = 0; kq < kSB / 4; ++kq) {
            const T bv = Ss[(4 * kq + lk) * kSR + lc];
#pragma unroll
            for (int t = 0; t < 2; ++t) y[t] = M::mma(dv[t][kq], bv, y[t]);
        }
        // y -= M_b Y_{b-1}
        if (b > 0) {
            T *ys = Ys[(b - 1) & 1];
            fetch_y(b - 1, ys);
            acc_t z[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) z[t] = acc_t{T(0), T(0), T(0), T(0)};
            if (m_in_b) {
                prod(ub, 0, ys, z);
                if (U == 2) prod(ua, KQ, ys, z);
            } else {
                prod(ua, 0, ys, z);
                if (U == 2) prod(ub, KQ, ys, z);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) y[t][r] -= z[t][r];
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + 32 * w + 16 * t + M::row(lane, r);
                if (row < n) store_coherent(Y + row * kSR + lc, y[t][r]);
            }
        stores_done();
        __syncthreads();
        if (threadIdx.x == 0) store_coherent(flags + b, epoch);
        // X is read by no other block: update it after the announcement
        if (X && lc < nrhs) {
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t row = r0 + 32 * w + 16 * t + M::row(lane, r);
                    if (row < n) X[(rev ? n - 1 - row : row) + lc * ldx] += (double)y[t][r];
                }
        }
    }
}

// M_b = L_bb⁻¹ L_{b,b−1} for every block b ≥ 1, column-major in Mt
// (npad × 128: row r0 + r, column c at Mt[r0 + r + c·npad]).  Grid
// (nblk − 1) × 4, each workgroup 32 columns of one M_b.
template <typename T>
__global__ __launch_bounds__(256) void diag_sub_kernel(int64_t n, const T *__restrict__ Dt,
                                                       const T *__restrict__ L, int64_t ldl,
                                                       T *__restrict__ Mt, int64_t npad)
{
    __shared__ T Ls[kSB][33];
    const int t = threadIdx.x;
    const int64_t b = blockIdx.x + 1;
    const int c0 = 32 * blockIdx.y;
    const int64_t r0 = b * kSB, cb = (b - 1) * kSB + c0;
    for (int e = t; e < kSB * 32; e += 256) {
        const int k = e % kSB, c = e / kSB;
        const int64_t row = r0 + k;
        Ls[k][c] = row < n ? L[row + (cb + c) * ldl] : T(0);
    }
    __syncthreads();
    const int r = t >> 1, ch = 16 * (t & 1);
    const T *D = Dt + b * (kSB * kSB) + (int64_t)r * kSB;
    T acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = T(0);
    for (int k = 0; k < kSB; ++k) {
        const T d = D[k];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += d * Ls[k][ch + c];
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) Mt[r0 + r + (int64_t)(c0 + ch + c) * npad] = acc[c];
}

// ---------------------------------------------------- symmetric residual --
constexpr int kVB = 64;    // rows of a block
constexpr int kRB = 8;     // blocks per rectangle side (512 rows)

// Symmetric product Y = A X over the lower triangle of A, each element of A
// read once.  Workgroup (I, J), J ≤ I, takes the rectangle of 64×64 blocks
// A_ij with i in rectangle row I, j in rectangle column J (and j ≤ i):
//   y_i += A_ij x_j                 summed into its row partial  Prow[J]
//   y_j += A_ijᵀ x_i   (j < i)      summed into its column partial Pcol[I]
// (diagonal blocks are mirrored from their lower half in LDS).  A row of
// rectangle row R thus has partials Prow[J] for J ≤ R and Pcol[I] for I ≥ R,
// each written by exactly one workgroup; symv_finish_kernel adds them in a
// fixed order, so the result does not depend on scheduling.  The next block
// is loaded into registers while the current one is multiplied.
__global__ __launch_bounds__(kST, 2) void symv_rect_kernel(int64_t n, const double *__restrict__ A,
                                                           int64_t lda,
                                                           const double *__restrict__ Xt,
                                                           double *__restrict__ Prow,
                                                           double *__restrict__ Pcol,
                                                           int64_t npad, int NR)
{
    typedef SMfma<double> M;
    constexpr int LDT = kVB + 1;
    __shared__ double Ts[kVB * LDT];           // A_ij column-major: (r, c) at [c·65 + r]
    __shared__ double xi[kVB * kSR], xj[kVB * kSR];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int lc = lane & 15, lk = lane >> 4;
    int rid = blockIdx.x, J = 0;
    while (rid >= NR - J) {
        rid -= NR - J;
        ++J;
    }
    const int I = J + rid;
    const int64_t nb = (n + kVB - 1) / kVB;
    const int64_t i0 = (int64_t)I * kRB, j0 = (int64_t)J * kRB;

    // thread t holds rows lane, columns w + 4u (u < 16) of a block
    auto gload = [&](int64_t i, int64_t j, double (&r)[16]) {
        const int64_t row = i * kVB + lane;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int64_t col = j * kVB + w + 4 * u;
            r[u] = (row < n && col < n) ? A[row + col * lda] : 0.0;
        }
    };
    d4 yr[kRB];
#pragma unroll
    for (int b = 0; b < kRB; ++b) yr[b] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[16];
    gload(i0, j0, ra);
    for (int bj = 0; bj < kRB; ++bj) {
        const int64_t j = j0 + bj;
        if (j >= nb) break;
        d4 yc = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int bi = 0; bi < kRB; ++bi) {
            const int64_t i = i0 + bi;
            if (i >= nb || i < j) continue;
#pragma unroll
            for (int u = 0; u < 16; ++u) Ts[(w + 4 * u) * LDT + lane] = ra[u];
#pragma unroll
            for (int e = t; e < kVB * kSR; e += kST) {
                xi[e] = Xt[i * kVB * kSR + e];
                xj[e] = Xt[j * kVB * kSR + e];
            }
            __syncthreads();
            if (i == j) {
                // upper half of a diagonal block from its lower half
                for (int e = t; e < kVB * kVB; e += kST) {
                    const int r = e % kVB, c = e / kVB;
                    if (r < c) Ts[c * LDT + r] = Ts[r * LDT + c];
                }
                __syncthreads();
            }
            // the next block of the rectangle: (i+1, j), else the first of column j+1
            {
                int64_t ni = i + 1, nj = j;
                if (bi + 1 >= kRB || ni >= nb) {
                    nj = j + 1;
                    ni = nj > i0 ? nj : i0;
                }
                if (nj < j0 + kRB && nj < nb && ni < nb && ni < i0 + kRB) gload(ni, nj, ra);
            }
#pragma unroll
            for (int kq = 0; kq < kVB / 4; ++kq)
                yr[bi] = M::mma(Ts[(4 * kq + lk) * LDT + 16 * w + lc], xj[(4 * kq + lk) * kSR + lc],
                                yr[bi]);
            if (i != j) {
#pragma unroll
                for (int kq = 0; kq < kVB / 4; ++kq)
                    yc = M::mma(Ts[(16 * w + lc) * LDT + 4 * kq + lk], xi[(4 * kq + lk) * kSR + lc],
                                yc);
            }
            __syncthreads();
        }
        double *pc = Pcol + ((int64_t)I * npad + j * kVB + 16 * w) * kSR;
#pragma unroll
        for (int r = 0; r < 4; ++r) pc[M::row(lane, r) * kSR + lc] = yc[r];
    }
#pragma unroll
    for (int bi = 0; bi < kRB; ++bi) {
        const int64_t i = i0 + bi;
        if (i >= nb) break;
        double *pr = Prow + ((int64_t)J * npad + i * kVB + 16 * w) * kSR;
#pragma unroll
        for (int r = 0; r < 4; ++r) pr[M::row(lane, r) * kSR + lc] = yr[bi][r];
    }
}

// R = B − A X with A X = Σ_{J≤R} Prow[J] + Σ_{I≥R} Pcol[I] for the rows of
// rectangle row R, added in that order (type T, row-major n×16);
// nrm[k] += ‖r_k‖², nrm[16+k] += ‖x_k‖² over this workgroup's 64 rows.
template <typename T>
__global__ __launch_bounds__(kST) void symv_finish_kernel(int64_t n, const double *__restrict__ Prow,
                                                          const double *__restrict__ Pcol,
                                                          int64_t npad, int NR,
                                                          const double *__restrict__ Xt,
                                                          const double *__restrict__ B,
                                                          int64_t ldb, int nrhs, T *__restrict__ R,
                                                          double *nrm)
{
    __shared__ double sr[2 * kST];
    const int64_t r0 = (int64_t)blockIdx.x * kVB;
    const int Rr = (int)(blockIdx.x / kRB);
    // every element this thread visits has k = threadIdx.x % 16 (256 is a
    // multiple of 16): accumulate locally, reduce over the 16 threads of each
    // k, one atomic per k and workgroup
    double rs = 0.0, xs = 0.0;
    for (int e = threadIdx.x; e < kVB * kSR; e += kST) {
        const int rl = e / kSR, k = e % kSR;
        const int64_t row = r0 + rl;
        if (row >= n) continue;
        double v = 0.0;
        if (k < nrhs) {
            double ax = 0.0;
            for (int q = 0; q <= Rr; ++q) ax += Prow[((int64_t)q * npad + row) * kSR + k];
            for (int q = Rr; q < NR; ++q) ax += Pcol[((int64_t)q * npad + row) * kSR + k];
            v = B[row + k * ldb] - ax;
            const double x = Xt[row * kSR + k];
            xs += x * x;
            rs += v * v;
        }
        R[row * kSR + k] = (T)v;
    }
    sr[threadIdx.x] = rs;
    sr[kST + threadIdx.x] = xs;
    __syncthreads();
    if (threadIdx.x < kSR && (int)threadIdx.x < nrhs) {
        double a = 0.0, c = 0.0;
        for (int q = threadIdx.x; q < kST; q += kSR) {
            a += sr[q];
            c += sr[kST + q];
        }
        atomicAdd(nrm + threadIdx.x, a);
        atomicAdd(nrm + kSR + threadIdx.x, c);
    }
}
// Xt (row-major npad×16) = X's first nrhs columns, zero elsewhere.
__global__ void rowmajor16_kernel(int64_t n, int64_t npad, int nrhs, const double *__restrict__ X,
                                  int64_t ldx, double *__restrict__ Xt)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < npad * kSR) {
        const int64_t row = e / kSR;
        const int k = (int)(e % kSR);
        Xt[e] = (k < nrhs && row < n) ? X[row + k * ldx] : 0.0;
    }
}
}  // namespace

template <typename T>
int trsv16(Handle *h, hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Dt, const T *Mt,
           const T *R, T *Y, int rev, double *X, int64_t ldx, int nrhs, int *flags, int epoch,
           int *err)
{
    if (n <= 0) return GLA_OK;
    if (nrhs > kSR) {
        set_error("trsv16: at most 16 right-hand sides per pass");
        return GLA_ERR_UNSUPPORTED;
    }
    const int64_t nblk = cdiv(n, kSB);
    const int grid = (int)std::min<int64_t>(nblk, h->num_cus);
    hipLaunchKernelGGL((trsv16_kernel<T>), dim3(grid), dim3(kTT), 0, s, n, nblk, L, ldl, Dt, Mt,
                       nblk * kSB, R, Y, rev, X, ldx, nrhs, flags, epoch, err);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template <typename T>
int diag_sub(hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Dt, T *Mt)
{
    const int64_t nblk = cdiv(n, kSB);
    if (nblk < 2) return GLA_OK;
    hipLaunchKernelGGL((diag_sub_kernel<T>), dim3((unsigned)(nblk - 1), 4), dim3(256), 0, s, n, Dt,
                       L, ldl, Mt, nblk * kSB);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// R = B − sym(A) X; Xt scratch: npad×16 FP64 followed by the rectangle
// partials (symv_scratch(n) doubles in all); nrm[32] must be zero on entry.
template <typename T>
int symv_residual(hipStream_t s, int64_t n, const double *A, int64_t lda, const double *X,
                  int64_t ldx, const double *B, int64_t ldb, int nrhs, double *Xt, T *R,
                  double *nrm)
{
    if (n <= 0) return GLA_OK;
    const int64_t nb = cdiv(n, kVB), npad = nb * kVB;
    const int NR = (int)cdiv(nb, kRB);
    double *Prow = Xt + npad * kSR;
    double *Pcol = Prow + (int64_t)NR * npad * kSR;
    hipLaunchKernelGGL(rowmajor16_kernel, dim3((unsigned)cdiv(npad * kSR, 256)), dim3(256), 0, s,
                       n, npad, nrhs, X, ldx, Xt);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(symv_rect_kernel, dim3((unsigned)(NR * (NR + 1) / 2)), dim3(kST), 0, s, n,
                       A, lda, Xt, Prow, Pcol, npad, NR);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((symv_finish_kernel<T>), dim3((unsigned)nb), dim3(kST), 0, s, n, Prow, Pcol,
                       npad, NR, Xt, B, ldb, nrhs, R, nrm);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

size_t symv_scratch(int64_t n)
{
    const int64_t nb = cdiv(n, kVB), NR = cdiv(nb, kRB);
    return (size_t)(nb * kVB) * kSR * (1 + 2 * NR);
}

template int trsv16<double>(Handle *, hipStream_t, int64_t, const double *, int64_t,
                            const double *, const double *, const double *, double *, int,
                            double *, int64_t, int, int *, int, int *);
template int trsv16<float>(Handle *, hipStream_t, int64_t, const float *, int64_t, const float *,
                           const float *, const float *, float *, int, double *, int64_t, int,
                           int *, int, int *);
template int diag_sub<double>(hipStream_t, int64_t, const double *, int64_t, const double *,
                              double *);
template int diag_sub<float>(hipStream_t, int64_t, const float *, int64_t, const float *, float *);
template int symv_residual<double>(hipStream_t, int64_t, const double *, int64_t, const double *,
                                   int64_t, const double *, int64_t, int, double *, double *,
                                   double *);
template int symv_residual<float>(hipStream_t, int64_t, const double *, int64_t, const double *,
                                  int64_t, const double *, int64_t, int, double *, float *,
                                  double *);

}  // namespace gla
