// Two-sided Householder reductions (FP64/FP32, real):
//   symtriLower!/symtriUpper!  src/eigenSelfAdjoint.jl:450-564  (A = Q T Qᵀ)
//   bidiagonalize!             src/svd.jl:328-381               (A = Ql B Qrᵀ)
//
// Both produce the reference's reflectors (Julia stdlib reflector!: β =
// −copysign(‖x‖, x₁), v₁ = 1, v = x/(x₁ − β), τ = (x₁ − β)/(−β)) in the
// reference's packing; the per-column rank-2 (or left/right) updates the
// reference applies at once are deferred across a panel of nb columns and
// applied as MFMA GEMMs at the panel's end — the same matrices up to
// rounding.
//
// symtri (lower; upper runs as lower on J A J, J the reversal, which maps
// symtriUpper!'s bottom-up sweep onto symtriLower!'s).  For column c of a
// panel starting at k (j = c − k), with the trailing matrix Ã kept as it
// was at the panel's start and the panel's earlier reflectors in V, W
// (Ã_now = Ã − V Wᵀ − W Vᵀ):
//   tri_col:    a = A[c:n, c] − V W[c,:]ᵀ − W V[c,:]ᵀ (the column brought up
//               to date; a[0] is the diagonal d_c); partial ‖a[1:]‖²;
//   tri_symv:   the reflector's scalars from those partials, then
//               p = Ã[c+1:n, c+1:n] v (the symmetric product from the lower
//               triangle, each 64×64 block read once and used for both
//               A_IJ v_J and A_IJᵀ v_I, partials per block in a fixed
//               order) and Wᵀv, Vᵀv partials;
//   tri_w:      v written into A and V, u = τ (p − V (Wᵀv) − W (Vᵀv)) into
//               W, partial vᵀu;
//   (next tri_col, or tri_wfin at the panel's end): w = u − (τ vᵀu / 2) v,
//               so that Ã_new = Ã_now − v wᵀ − w vᵀ = H Ã_now H.
// Panel end: Ã[k+nb:, k+nb:] −= V Wᵀ + W Vᵀ (two lower-triangle GEMMs).
#include <algorithm>

#include "gla_internal.hpp"

namespace gla {

namespace {

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

constexpr int kTB = 64;       // symmetric-product block
constexpr int kPanel = 32;    // panel width (columns of V, W)

// Scalars of the current column, written by tri_symv's first workgroup.
enum { S_TAU = 0, S_XI = 1, S_BETA = 2, S_UNEXT = 3, S_COUNT = 8 };

constexpr int kRPB = 64;      // rows per workgroup of the row kernels (4 lanes per row)

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    return wave_sum_dpp(v);
}

// Fixed-order sum of part[0..cnt) by one workgroup (every thread gets it).
template <typename T>
__device__ T block_sum_fixed(const T *part, int cnt, T *sh)
{
    T s = T(0);
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) s += part[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    T r = T(0);
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
    __syncthreads();
    return r;
}

// The reflector! scalars of column c from ‖a[1:]‖² (partials) and a[1]:
// ξ₁ = a₁ + ν (the divisor of the tail), τ = ξ₁/ν, β = −ν; a zero norm
// leaves the column alone (τ = 0, ξ₁ = 1).
template <typename T>
__device__ __forceinline__ void reflector_scalars(T nrm2, T x1, T &xi, T &tau, T &beta)
{
    const T normu = sqrt(nrm2);
    if (normu == T(0)) {
        xi = T(1);
        tau = T(0);
        beta = x1;
        return;
    }
    const T nu = copysign(normu, x1);
    xi = x1 + nu;
    tau = xi / nu;
    beta = -nu;
}

// Column c brought up to date with the panel's reflectors 0..j−1; before
// that, W[:, j−1] (holding u) becomes w = u − (ξ/2) v for rows ≥ c.  Four
// lanes per row, lane k taking the reflectors l ≡ k (mod 4) (loads from
// clamped indices, all issued at once), summed over the quad.
template <typename T>
__global__ __launch_bounds__(256) void tri_col_kernel(int64_t n, int64_t c, int j, T *A,
                                                      int64_t lda, T *V, T *W, int64_t ldv,
                                                      const T *qpart, int qcnt, const T *scal,
                                                      T *npart)
{
    __shared__ T sh[8];
    __shared__ T wc[kPanel], vc[kPanel];
    const int t = threadIdx.x, k = t & 3;
    const int64_t i = c + (int64_t)blockIdx.x * kRPB + (t >> 2);
    const bool in = i < n;
    const int64_t ic = in ? i : n - 1;
    T wprev_c = T(0), half_xi = T(0);
    if (j > 0) {
        const T vu = block_sum_fixed(qpart, qcnt, sh);
        half_xi = T(0.5) * scal[S_TAU] * vu;              // ξ/2 = τ vᵀu / 2
        wprev_c = scal[S_UNEXT] - half_xi;                // row c: v_c = 1
    }
    if (t < kPanel) {
        wc[t] = t < j ? (t == j - 1 ? wprev_c : W[c + (int64_t)t * ldv]) : T(0);
        vc[t] = t < j ? V[c + (int64_t)t * ldv] : T(0);
    }
    __syncthreads();
    T part = T(0);
    if (j > 0) {
        T wv[kPanel / 4], vv[kPanel / 4];
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = k + 4 * u, lc = l < j ? l : 0;
            wv[u] = W[ic + (int64_t)lc * ldv];
            vv[u] = V[ic + (int64_t)lc * ldv];
        }
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = k + 4 * u;
            T wil = wv[u];
            if (l == j - 1) {
                wil -= half_xi * vv[u];
                if (in) W[i + (int64_t)l * ldv] = wil;
            }
            if (l < j) part += vv[u] * wc[l] + wil * vc[l];
        }
    }
    part = quad_sum(part);
    T a = T(0);
    if (in) {
        a = A[i + c * lda] - part;
        if (k == 0) A[i + c * lda] = a;
    }
    T s = (in && i > c && k == 0) ? a * a : T(0);
    s = wave_sum(s);
    if ((t & 63) == 0) sh[t >> 6] = s;
    __syncthreads();
    if (t == 0) npart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// After the panel's last column: W[:, j] (u) becomes w for rows ≥ c.
template <typename T>
__global__ void tri_wfin_kernel(int64_t n, int64_t c, int j, T *V, T *W, int64_t ldv,
                                const T *qpart, int qcnt, const T *scal)
{
    __shared__ T sh[8];
    const T vu = block_sum_fixed(qpart, qcnt, sh);
    const T half_xi = T(0.5) * scal[S_TAU] * vu;
    const int64_t i = c + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) W[i + (int64_t)j * ldv] -= half_xi * V[i + (int64_t)j * ldv];
}

// v_r of row r ≥ r0 = c + 1 from the unscaled column.
template <typename T>
__device__ __forceinline__ T vval(const T *A, int64_t lda, int64_t c, int64_t r, T xi)
{
    return r == c + 1 ? T(1) : A[r + c * lda] / xi;
}

// p = Ã v over the lower-stored trailing block, blocks (I, J), I ≥ J, of
// 64×64; P[s·m + r] is the contribution of the block pairing row block
// r/64 with block s (rows relative to r0).  Diagonal blocks add Wᵀv and
// Vᵀv partials Y1[I·kPanel + l], Y2[I·kPanel + l].
template <typename T>
__global__ __launch_bounds__(256) void tri_symv_kernel(int64_t n, int64_t c, int j, const T *A,
                                                       int64_t lda, const T *V, const T *W,
                                                       int64_t ldv, const T *npart, int ncnt,
                                                       T *P, T *Y1, T *Y2, T *scal, T *tau_out)
{
    __shared__ T sh[8];
    __shared__ T srow[4][kTB];
    const int64_t r0 = c + 1, m = n - r0;
    // block (I, J) from the linear id: row I, J ≤ I
    const int64_t q = blockIdx.x;
    int64_t I = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) / 2.0);
    while ((I + 1) * (I + 2) / 2 <= q) ++I;
    while (I * (I + 1) / 2 > q) --I;
    const int64_t J = q - I * (I + 1) / 2;
    const T nrm2 = block_sum_fixed(npart, ncnt, sh);
    T xi, tau, beta;
    reflector_scalars(nrm2, A[r0 + c * lda], xi, tau, beta);
    if (q == 0 && threadIdx.x == 0) {
        scal[S_TAU] = tau;
        scal[S_XI] = xi;
        scal[S_BETA] = beta;
        tau_out[c] = tau;
    }
    const int t = threadIdx.x, a = t & 15, b = t >> 4, lane = t & 63, wv = t >> 6;
    // this thread's rows RB + 16u (consecutive lanes read consecutive rows,
    // so each load instruction reads whole cache lines) and columns C0 + u;
    // the loads are unconditional (clamped indices), out-of-range values
    // and the upper part of a diagonal block replaced by zero
    const int64_t RB = r0 + I * kTB + a, C0 = r0 + J * kTB + 4 * b;
    T x[4][4];
    T vr[4], vcv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t rr = RB + 16 * u, cc = C0 + u;
        vr[u] = rr < n ? vval(A, lda, c, rr, xi) : T(0);
        vcv[u] = cc < n ? vval(A, lda, c, cc, xi) : T(0);
    }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int64_t gi = RB + 16 * rr, gj = C0 + cc;
            const int64_t gic = gi < n ? gi : n - 1, gjc = gj < n ? gj : n - 1;
            const T v = A[gic + gjc * lda];
            x[cc][rr] = (gi < n && gj < n && gi >= gj) ? v : T(0);
        }
    // row sums (A_IJ v_J) and column sums (A_IJᵀ v_I, strictly lower part
    // only on a diagonal block)
    T s[4], qc[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        T acc = T(0);
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) acc += x[cc][rr] * vcv[cc];
        s[rr] = acc;
    }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        T acc = T(0);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const bool use = I != J || (RB + 16 * rr > C0 + cc);
            acc += use ? x[cc][rr] * vr[rr] : T(0);
        }
        qc[cc] = acc;
    }
    // s: reduce over b (lane bits 4, 5, then the four waves)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        s[rr] += __shfl_xor(s[rr], 16);
        s[rr] += __shfl_xor(s[rr], 32);
    }
    if (lane < 16)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) srow[wv][a + 16 * rr] = s[rr];
    // qc: reduce over a (lane bits 0..3)
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) qc[cc] = row_sum16(qc[cc]);
    __syncthreads();
    if (I == J) {
        // one slot for the diagonal block: row sums plus the transposed part
        __shared__ T scol[kTB];
        if (a == 0)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) scol[4 * b + cc] = qc[cc];
        __syncthreads();
        if (t < kTB) {
            const int64_t r = I * kTB + t;
            if (r < m) P[J * m + r] = srow[0][t] + srow[1][t] + srow[2][t] + srow[3][t] + scol[t];
        }
        // Wᵀv and Vᵀv over this block's rows: wave wv takes l ≡ wv (mod 4),
        // all loads issued before the sums
        const int64_t gr = r0 + I * kTB + lane;
        const int64_t grc = gr < n ? gr : n - 1;
        const T vg = gr < n ? vval(A, lda, c, gr, xi) : T(0);
        T wl[kPanel / 4], vl[kPanel / 4];
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = wv + 4 * u, lc = l < j ? l : 0;
            wl[u] = W[grc + (int64_t)lc * ldv];
            vl[u] = V[grc + (int64_t)lc * ldv];
        }
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = wv + 4 * u;
            if (l < j) {
                const T y1 = wave_sum(wl[u] * vg);
                const T y2 = wave_sum(vl[u] * vg);
                if (lane == 0) {
                    Y1[I * kPanel + l] = y1;
                    Y2[I * kPanel + l] = y2;
                }
            }
        }
    } else {
        if (t < kTB) {
            const int64_t r = I * kTB + t;
            if (r < m) P[J * m + r] = srow[0][t] + srow[1][t] + srow[2][t] + srow[3][t];
        }
        if (a == 0)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                const int64_t r = J * kTB + 4 * b + cc;
                if (r < m) P[I * m + r] = qc[cc];
            }
    }
}

// v into A (below the subdiagonal) and V, β on the subdiagonal;
// u = τ (p − V y1 − W y2) into W[:, j]; partial vᵀu.  Four lanes per row:
// lane k sums the block contributions s ≡ k (mod 4) of p and the
// reflectors l ≡ k (mod 4), then the quad sums; y1, y2 (Wᵀv, Vᵀv) are
// summed over the blocks' partials by the whole workgroup, fixed order.
template <typename T>
__global__ __launch_bounds__(256) void tri_w_kernel(int64_t n, int64_t c, int j, T *A, int64_t lda,
                                                    T *V, T *W, int64_t ldv, const T *P, int nbk,
                                                    const T *Y1, const T *Y2, T *scal, T *qpart)
{
    __shared__ T sh[8];
    __shared__ T y1[kPanel], y2[kPanel];
    __shared__ T ys[2][8][kPanel];
    const int t = threadIdx.x, k = t & 3;
    const int64_t r0 = c + 1, m = n - r0;
    const T tau = scal[S_TAU], xi = scal[S_XI], beta = scal[S_BETA];
    {
        const int l = t & (kPanel - 1), pp = t / kPanel;   // 8 parts of the block sums
        T s1 = T(0), s2 = T(0);
        if (l < j)
            for (int I = pp; I < nbk; I += 8) {
                s1 += Y1[I * kPanel + l];
                s2 += Y2[I * kPanel + l];
            }
        ys[0][pp][l] = s1;
        ys[1][pp][l] = s2;
    }
    __syncthreads();
    if (t < kPanel) {
        T s1 = T(0), s2 = T(0);
#pragma unroll
        for (int pp = 0; pp < 8; ++pp) {
            s1 += ys[0][pp][t];
            s2 += ys[1][pp][t];
        }
        y1[t] = s1;
        y2[t] = s2;
    }
    __syncthreads();
    const int64_t r = r0 + (int64_t)blockIdx.x * kRPB + (t >> 2);
    const bool in = r < n;
    const int64_t rc = in ? r : n - 1, rr = rc - r0;
    T p0 = T(0), p1 = T(0);
    int sblk = k;
    for (; sblk + 4 < nbk; sblk += 8) {
        p0 += P[(int64_t)sblk * m + rr];
        p1 += P[(int64_t)(sblk + 4) * m + rr];
    }
    if (sblk < nbk) p0 += P[(int64_t)sblk * m + rr];
    T p = p0 + p1;
    if (j > 0) {
        T wv[kPanel / 4], vv[kPanel / 4];
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = k + 4 * u, lc = l < j ? l : 0;
            vv[u] = V[rc + (int64_t)lc * ldv];
            wv[u] = W[rc + (int64_t)lc * ldv];
        }
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = k + 4 * u;
            if (l < j) p -= vv[u] * y1[l] + wv[u] * y2[l];
        }
    }
    p = quad_sum(p);
    T vu = T(0);
    if (in && k == 0) {
        const T v = r == r0 ? T(1) : A[r + c * lda] / xi;
        const T u = tau * p;
        W[r + (int64_t)j * ldv] = u;
        V[r + (int64_t)j * ldv] = v;
        A[r + c * lda] = r == r0 ? beta : v;
        if (r == r0) scal[S_UNEXT] = u;
        vu = v * u;
    }
    vu = wave_sum(vu);
    if ((t & 63) == 0) sh[t >> 6] = vu;
    __syncthreads();
    if (t == 0) qpart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// A[i, j] <-> A[n−1−i, n−1−j] over the whole matrix (J A J in place).
template <typename T>
__global__ void reverse_kernel(int64_t n, T *A, int64_t lda)
{
    const int64_t i = (int64_t)blockIdx.x * 16 + threadIdx.x;
    const int64_t j = (int64_t)blockIdx.y * 16 + threadIdx.y;
    if (i >= n || j >= n) return;
    const int64_t i2 = n - 1 - i, j2 = n - 1 - j;
    // each pair once: the element with the smaller column-major index swaps
    if (i + j * n < i2 + j2 * n) {
        const T a = A[i + j * lda];
        A[i + j * lda] = A[i2 + j2 * lda];
        A[i2 + j2 * lda] = a;
    }
}

template <typename T>
__global__ void zero_kernel(int64_t cnt, T *p)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < cnt) p[i] = T(0);
}


// ------------------------------------------------------ bidiagonalize! ----
// Tall case (m ≥ n), panel [k, k+w), column i = k + j.  The trailing block
// Ã keeps its panel-start values wherever a GEMV reads it (the column and
// row brought up to date lie outside those regions); V, X (m rows) and
// Y, U (n rows) hold the panel's left vectors, A u products, Aᵀ v products
// and right vectors: A_now = Ã − V Yᵀ − X Uᵀ.
enum { B_LTAU = 0, B_LXI = 1, B_LBETA = 2, B_RTAU = 3, B_RXI = 4, B_RBETA = 5 };

// ‖A[i:m, i]‖² partials (the panel's first column, nothing to apply).
template <typename T>
__global__ void bd_col0_kernel(int64_t m, int64_t i, const T *A, int64_t lda, T *npart)
{
    __shared__ T sh[8];
    const int64_t r = i + (int64_t)blockIdx.x * 256 + threadIdx.x;
    T a = r < m ? A[r + i * lda] : T(0);
    T sq = wave_sum(a * a);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) npart[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Left reflector scalars of column i, then P[rb·C + c'] = Σ over the 64-row
// block rb of Ã[r, i+1+c'] v_r (Ãᵀ v), and on cb = 0 blocks the partials
// Cv1[rb·kPanel + l] = Σ V[r,l] v_r, Cv2 = Σ X[r,l] v_r (l < j).
template <typename T>
__global__ __launch_bounds__(256) void bd_gemvt_kernel(int64_t m, int64_t n, int64_t i, int j,
                                                       const T *A, int64_t lda, const T *V,
                                                       const T *X, int64_t ldm,
                                                       const T *npart, int ncnt, int64_t ncb,
                                                       T *P, T *Cv1, T *Cv2, T *scal, T *taul)
{
    __shared__ T sh[8];
    const T nrm2 = block_sum_fixed(npart, ncnt, sh);
    T xi, tau, beta;
    reflector_scalars(nrm2, A[i + i * lda], xi, tau, beta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[B_LTAU] = tau;
        scal[B_LXI] = xi;
        scal[B_LBETA] = beta;
        taul[i] = tau;
    }
    const int64_t C = n - i - 1;
    if (C <= 0) return;
    const int64_t rb = blockIdx.x / ncb, cb = blockIdx.x % ncb;
    const int t = threadIdx.x, a = t & 15, b = t >> 4, lane = t & 63, wv = t >> 6;
    // rows RB + 16u (whole cache lines per load instruction), columns
    // C0 + cc; unconditional loads from clamped indices
    const int64_t RB = i + rb * kTB + a, C0 = i + 1 + cb * kTB + 4 * b;
    T vr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t r = RB + 16 * u;
        vr[u] = r < m ? (r == i ? T(1) : A[r + i * lda] / xi) : T(0);
    }
    T x[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int64_t gr = RB + 16 * rr, gc = C0 + cc;
            const int64_t grc = gr < m ? gr : m - 1, gcc = gc < n ? gc : n - 1;
            const T v = A[grc + gcc * lda];
            x[cc][rr] = (gr < m && gc < n) ? v : T(0);
        }
    T qc[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        T acc = T(0);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) acc += x[cc][rr] * vr[rr];
        qc[cc] = acc;
    }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) qc[cc] = row_sum16(qc[cc]);
    if (a == 0)
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            const int64_t c = cb * kTB + 4 * b + cc;
            if (c < C) P[rb * C + c] = qc[cc];
        }
    if (cb == 0) {
        const int64_t gr = i + rb * kTB + lane;
        const int64_t grc = gr < m ? gr : m - 1;
        const T vg = gr < m ? (gr == i ? T(1) : A[gr + i * lda] / xi) : T(0);
        T a1[kPanel / 4], a2[kPanel / 4];
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = wv + 4 * u, lc = l < j ? l : 0;
            a1[u] = V[grc + (int64_t)lc * ldm];
            a2[u] = X[grc + (int64_t)lc * ldm];
        }
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = wv + 4 * u;
            if (l < j) {
                const T c1 = wave_sum(a1[u] * vg);
                const T c2 = wave_sum(a2[u] * vg);
                if (lane == 0) {
                    Cv1[rb * kPanel + l] = c1;
                    Cv2[rb * kPanel + l] = c2;
                }
            }
        }
    }
}

// Sum of the per-block partials part[b·kPanel + l] over b < nb for l < lmax,
// by the whole workgroup in a fixed order (eight interleaved parts, then
// the parts in order); the result in out[l] (zero for l ≥ lmax).
template <typename T>
__device__ __forceinline__ void panel_partials(const T *part, int64_t nb, int lmax, T *out,
                                               T (*cs)[kPanel])
{
    const int t = threadIdx.x, l = t & (kPanel - 1), pp = t / kPanel;
    T s = T(0);
    if (l < lmax)
        for (int64_t b = pp; b < nb; b += 8) s += part[b * kPanel + l];
    cs[pp][l] = s;
    __syncthreads();
    if (t < kPanel) {
        T r = T(0);
#pragma unroll
        for (int q = 0; q < 8; ++q) r += cs[q][t];
        out[t] = r;
    }
}

// Item it of the workgroup (64 items, four lanes each) covers column
// c = i+1+it (y_c = τ (Ãᵀv − Y c1 − U c2)_c, row i brought up to date,
// partial ‖row‖²; lane k takes the row-block partials rb ≡ k and the
// reflectors l ≡ k (mod 4), summed over the quad) and row r = i+it (v into
// A and V).
template <typename T>
__global__ __launch_bounds__(256) void bd_y_kernel(int64_t m, int64_t n, int64_t i, int j, T *A,
                                                   int64_t lda, T *V, const T *X, int64_t ldm,
                                                   T *Y, const T *U, int64_t ldn, const T *P,
                                                   int64_t nrb, const T *Cv1, const T *Cv2,
                                                   const T *scal, T *npart2)
{
    __shared__ T sh[8];
    __shared__ T c1[kPanel], c2[kPanel], vi[kPanel], xi_[kPanel];
    __shared__ T cs[8][kPanel];
    const int t = threadIdx.x, k = t & 3;
    const T tau = scal[B_LTAU], xi = scal[B_LXI], beta = scal[B_LBETA];
    if (t < kPanel) {
        vi[t] = t < j ? V[i + (int64_t)t * ldm] : T(0);
        xi_[t] = t < j ? X[i + (int64_t)t * ldm] : T(0);
    }
    panel_partials(Cv1, nrb, j, c1, cs);
    __syncthreads();
    panel_partials(Cv2, nrb, j, c2, cs);
    __syncthreads();
    const int64_t it = (int64_t)blockIdx.x * kRPB + (t >> 2);
    const int64_t C = n - i - 1;
    T sq = T(0);
    if (C > 0) {
        const bool in = it < C;
        const int64_t tc = in ? it : C - 1, c = i + 1 + tc;
        T p0 = T(0), p1 = T(0);
        int64_t rb = k;
        for (; rb + 4 < nrb; rb += 8) {
            p0 += P[rb * C + tc];
            p1 += P[(rb + 4) * C + tc];
        }
        if (rb < nrb) p0 += P[rb * C + tc];
        T p = p0 + p1, rp = T(0);
        if (j > 0) {
            T yv[kPanel / 4], uv[kPanel / 4];
#pragma unroll
            for (int u = 0; u < kPanel / 4; ++u) {
                const int l = k + 4 * u, lc = l < j ? l : 0;
                yv[u] = Y[c + (int64_t)lc * ldn];
                uv[u] = U[c + (int64_t)lc * ldn];
            }
#pragma unroll
            for (int u = 0; u < kPanel / 4; ++u) {
                const int l = k + 4 * u;
                if (l < j) {
                    p -= yv[u] * c1[l] + uv[u] * c2[l];
                    rp += vi[l] * yv[u] + xi_[l] * uv[u];
                }
            }
        }
        p = quad_sum(p);
        rp = quad_sum(rp);
        if (in && k == 0) {
            const T y = tau * p;
            Y[c + (int64_t)j * ldn] = y;
            const T r = A[i + c * lda] - y - rp;
            A[i + c * lda] = r;
            sq = r * r;
        }
    }
    const int64_t r = i + it;
    if (k == 0 && r < m) {
        const T v = r == i ? T(1) : A[r + i * lda] / xi;
        V[r + (int64_t)j * ldm] = v;
        A[r + i * lda] = r == i ? beta : v;
    }
    sq = wave_sum(sq);
    if ((t & 63) == 0) sh[t >> 6] = sq;
    __syncthreads();
    if (t == 0) npart2[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Right reflector scalars of row i, then P[cb·R + r'] = Σ over the 64-column
// block cb of Ã[i+1+r', c] u_c (Ã u), and on rb = 0 blocks
// Cu1[cb·kPanel + l] = Σ Y[c,l] u_c (l ≤ j), Cu2 = Σ U[c,l] u_c (l < j).
template <typename T>
__global__ __launch_bounds__(256) void bd_gemvn_kernel(int64_t m, int64_t n, int64_t i, int j,
                                                       const T *A, int64_t lda, const T *Y,
                                                       const T *U, int64_t ldn,
                                                       const T *npart2, int ncnt, int64_t ncb,
                                                       T *P, T *Cu1, T *Cu2, T *scal, T *taur)
{
    __shared__ T sh[8];
    __shared__ T srow[4][kTB];
    const T nrm2 = block_sum_fixed(npart2, ncnt, sh);
    T xi, tau, beta;
    reflector_scalars(nrm2, A[i + (i + 1) * lda], xi, tau, beta);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[B_RTAU] = tau;
        scal[B_RXI] = xi;
        scal[B_RBETA] = beta;
        taur[i] = tau;
    }
    const int64_t R = m - i - 1;
    const int64_t rb = blockIdx.x / ncb, cb = blockIdx.x % ncb;
    const int t = threadIdx.x, a = t & 15, b = t >> 4, lane = t & 63, wv = t >> 6;
    // rows RB + 16u (whole cache lines per load instruction), columns
    // C0 + cc; unconditional loads from clamped indices
    const int64_t RB = i + 1 + rb * kTB + a, C0 = i + 1 + cb * kTB + 4 * b;
    T uc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t c = C0 + u;
        uc[u] = c < n ? (c == i + 1 ? T(1) : A[i + c * lda] / xi) : T(0);
    }
    T x[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int64_t gr = RB + 16 * rr, gc = C0 + cc;
            const int64_t grc = gr < m ? gr : m - 1, gcc = gc < n ? gc : n - 1;
            const T v = A[grc + gcc * lda];
            x[cc][rr] = (gr < m && gc < n) ? v : T(0);
        }
    T s[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        T acc = T(0);
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) acc += x[cc][rr] * uc[cc];
        s[rr] = acc;
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        s[rr] += __shfl_xor(s[rr], 16);
        s[rr] += __shfl_xor(s[rr], 32);
    }
    if (lane < 16)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) srow[wv][a + 16 * rr] = s[rr];
    __syncthreads();
    if (t < kTB) {
        const int64_t r = rb * kTB + t;
        if (r < R) P[cb * R + r] = srow[0][t] + srow[1][t] + srow[2][t] + srow[3][t];
    }
    if (rb == 0) {
        const int64_t gc = i + 1 + cb * kTB + lane;
        const int64_t gcc = gc < n ? gc : n - 1;
        const T ug = gc < n ? (gc == i + 1 ? T(1) : A[i + gc * lda] / xi) : T(0);
        T a1[kPanel / 4], a2[kPanel / 4];
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = wv + 4 * u, l1 = l <= j ? l : 0, l2 = l < j ? l : 0;
            a1[u] = Y[gcc + (int64_t)l1 * ldn];
            a2[u] = U[gcc + (int64_t)l2 * ldn];
        }
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = wv + 4 * u;
            if (l <= j) {
                const T c1 = wave_sum(a1[u] * ug);
                const T c2 = wave_sum(l < j ? a2[u] * ug : T(0));
                if (lane == 0) {
                    Cu1[cb * kPanel + l] = c1;
                    Cu2[cb * kPanel + l] = c2;
                }
            }
        }
    }
}

// Item it (64 per workgroup, four lanes each) covers row r = i+1+it
// (x_r = τ (Ã u − V c3 − X c4)_r; then, unless `last`, column i+1 brought
// up to date with reflectors 0..j and its partial ‖·‖²; lane k takes the
// column-block partials cb ≡ k and the reflectors l ≡ k (mod 4), summed over
// the quad) and column c = i+1+it (u into A and U).
template <typename T>
__global__ __launch_bounds__(256) void bd_x_kernel(int64_t m, int64_t n, int64_t i, int j,
                                                   int last, T *A, int64_t lda, const T *V, T *X,
                                                   int64_t ldm, const T *Y, T *U, int64_t ldn,
                                                   const T *P, int64_t ncb, const T *Cu1,
                                                   const T *Cu2, const T *scal, T *npart)
{
    __shared__ T sh[8];
    __shared__ T c3[kPanel], c4[kPanel], yn[kPanel], un[kPanel];
    __shared__ T cs[8][kPanel];
    const int t = threadIdx.x, k = t & 3;
    const T tau = scal[B_RTAU], xi = scal[B_RXI], beta = scal[B_RBETA];
    const int64_t i1 = i + 1;
    if (t < kPanel) {
        yn[t] = t <= j ? Y[i1 + (int64_t)t * ldn] : T(0);
        un[t] = t <= j ? (t == j ? T(1) : U[i1 + (int64_t)t * ldn]) : T(0);
    }
    panel_partials(Cu1, ncb, j + 1, c3, cs);
    __syncthreads();
    panel_partials(Cu2, ncb, j, c4, cs);
    __syncthreads();
    const int64_t it = (int64_t)blockIdx.x * kRPB + (t >> 2);
    const int64_t R = m - i - 1;
    T sq = T(0);
    if (R > 0) {
        const bool in = it < R;
        const int64_t tr = in ? it : R - 1, r = i1 + tr;
        T q0 = T(0), q1 = T(0);
        int64_t cb = k;
        for (; cb + 4 < ncb; cb += 8) {
            q0 += P[cb * R + tr];
            q1 += P[(cb + 4) * R + tr];
        }
        if (cb < ncb) q0 += P[cb * R + tr];
        T q = q0 + q1, ap = T(0);
        T vv[kPanel / 4], xv[kPanel / 4];
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = k + 4 * u, lv = l <= j ? l : 0, lx = l < j ? l : 0;
            vv[u] = V[r + (int64_t)lv * ldm];
            xv[u] = X[r + (int64_t)lx * ldm];
        }
#pragma unroll
        for (int u = 0; u < kPanel / 4; ++u) {
            const int l = k + 4 * u;
            if (l <= j) q -= vv[u] * c3[l];
            if (l < j) {
                q -= xv[u] * c4[l];
                ap += vv[u] * yn[l] + xv[u] * un[l];
            }
        }
        q = quad_sum(q);
        ap = quad_sum(ap);
        if (in && k == 0) {
            const T x = tau * q;
            X[r + (int64_t)j * ldm] = x;
            if (!last && i1 < n) {
                // the l = j term: V[r][j] y_n[j] + x u_n[j]
                const T a = A[r + i1 * lda] - ap - (V[r + (int64_t)j * ldm] * yn[j] + x * un[j]);
                A[r + i1 * lda] = a;
                sq = a * a;
            }
        }
    }
    const int64_t c = i1 + it;
    if (k == 0 && c < n) {
        const T u = c == i1 ? T(1) : A[i + c * lda] / xi;
        U[c + (int64_t)j * ldn] = u;
        A[i + c * lda] = c == i1 ? beta : u;
    }
    sq = wave_sum(sq);
    if ((t & 63) == 0) sh[t >> 6] = sq;
    __syncthreads();
    if (t == 0) npart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// B (n×m, ld ldb) = Aᵀ for A m×n.
template <typename T>
__global__ void transpose_kernel(int64_t m, int64_t n, const T *A, int64_t lda, T *B, int64_t ldb)
{
    __shared__ T tile[16][17];
    const int64_t r0 = (int64_t)blockIdx.x * 16, c0 = (int64_t)blockIdx.y * 16;
    const int tx = threadIdx.x, ty = threadIdx.y;
    if (r0 + tx < m && c0 + ty < n) tile[ty][tx] = A[(r0 + tx) + (c0 + ty) * lda];
    __syncthreads();
    if (c0 + tx < n && r0 + ty < m) B[(c0 + tx) + (r0 + ty) * ldb] = tile[tx][ty];
}

}  // namespace

// Workspace elements of symtri for order n.
size_t symtri_scratch(int64_t n)
{
    const int64_t ldv = (n + 3) / 4 * 4;
    const int64_t nbk = cdiv(n, kTB);
    const int64_t g = cdiv(n, kRPB);
    return 2 * (size_t)ldv * kPanel + (size_t)nbk * n + 2 * (size_t)nbk * kPanel + 2 * g + 64;
}

// symtriLower! on the lower triangle of A (n×n); τ (n − 1 entries, device).
template <typename T>
int symtri_lower_dev(hipStream_t s, int64_t n, T *A, int64_t lda, T *tau, T *ws)
{
    if (n < 1) return GLA_OK;
    if (n > 1) {
        hipLaunchKernelGGL((zero_kernel<T>), dim3((unsigned)cdiv(n - 1, 256)), dim3(256), 0, s,
                           n - 1, tau);
        GLA_LAUNCH_CHECK();
    }
    if (n <= 2) return GLA_OK;
    const int64_t ldv = (n + 3) / 4 * 4;
    const int64_t nbk_max = cdiv(n, kTB);
    const int64_t gmax = cdiv(n, kRPB);
    T *V = ws, *W = V + ldv * kPanel, *P = W + ldv * kPanel, *Y1 = P + nbk_max * n,
      *Y2 = Y1 + nbk_max * kPanel, *npart = Y2 + nbk_max * kPanel, *qpart = npart + gmax,
      *scal = qpart + gmax;
    const int64_t ncols = n - 2;   // columns with a reflector (:462)
    for (int64_t k = 0; k < ncols; k += kPanel) {
        const int w = (int)std::min<int64_t>(kPanel, ncols - k);
        for (int j = 0; j < w; ++j) {
            const int64_t c = k + j;
            const int64_t m = n - c - 1;
            const unsigned g1 = (unsigned)cdiv(n - c, kRPB);
            const int qcnt = (int)cdiv(n - c, kRPB);   // the previous column's tri_w grid
            hipLaunchKernelGGL((tri_col_kernel<T>), dim3(g1), dim3(256), 0, s, n, c, j, A, lda, V,
                               W, ldv, qpart, qcnt, scal, npart);
            GLA_LAUNCH_CHECK();
            const int64_t nbk = cdiv(m, kTB);
            hipLaunchKernelGGL((tri_symv_kernel<T>), dim3((unsigned)(nbk * (nbk + 1) / 2)),
                               dim3(256), 0, s, n, c, j, (const T *)A, lda, (const T *)V,
                               (const T *)W, ldv, (const T *)npart, (int)g1, P, Y1, Y2, scal, tau);
            GLA_LAUNCH_CHECK();
            const unsigned g4 = (unsigned)cdiv(m, kRPB);
            hipLaunchKernelGGL((tri_w_kernel<T>), dim3(g4), dim3(256), 0, s, n, c, j, A, lda, V, W,
                               ldv, (const T *)P, (int)nbk, (const T *)Y1, (const T *)Y2, scal,
                               qpart);
            GLA_LAUNCH_CHECK();
        }
        // W's last column: u -> w over rows ≥ k + w
        const int64_t c1 = k + w;
        hipLaunchKernelGGL((tri_wfin_kernel<T>), dim3((unsigned)cdiv(n - c1, 256)), dim3(256), 0,
                           s, n, c1, w - 1, V, W, ldv, (const T *)qpart, (int)cdiv(n - c1, kRPB),
                           (const T *)scal);
        GLA_LAUNCH_CHECK();
        // Ã[c1:, c1:] −= V Wᵀ + W Vᵀ (lower triangle)
        const int64_t mt = n - c1;
        if (mt > 0) {
            GLA_TRY(gemm<T>(s, 'N', 'T', 1, mt, mt, w, T(-1), V + c1, ldv, W + c1, ldv, T(1),
                            A + c1 + c1 * lda, lda, 0));
            GLA_TRY(gemm<T>(s, 'N', 'T', 1, mt, mt, w, T(-1), W + c1, ldv, V + c1, ldv, T(1),
                            A + c1 + c1 * lda, lda, 0));
        }
    }
    return GLA_OK;
}

template <typename T>
int symtri_dev(hipStream_t s, char uplo, int64_t n, T *A, int64_t lda, T *tau, T *ws)
{
    const dim3 rg((unsigned)cdiv(n, 16), (unsigned)cdiv(n, 16));
    if (uplo == 'U') {
        hipLaunchKernelGGL((reverse_kernel<T>), rg, dim3(16, 16), 0, s, n, A, lda);
        GLA_LAUNCH_CHECK();
    }
    GLA_TRY(symtri_lower_dev<T>(s, n, A, lda, tau, ws));
    if (uplo == 'U') {
        hipLaunchKernelGGL((reverse_kernel<T>), rg, dim3(16, 16), 0, s, n, A, lda);
        GLA_LAUNCH_CHECK();
    }
    return GLA_OK;
}

template int symtri_dev<double>(hipStream_t, char, int64_t, double *, int64_t, double *, double *);
template int symtri_dev<float>(hipStream_t, char, int64_t, float *, int64_t, float *, float *);

// Workspace elements of bidiagonalize! for m×n.
size_t bidiag_scratch(int64_t m, int64_t n)
{
    const int64_t M = std::max(m, n), N = std::min(m, n);   // the tall problem
    const int64_t ldm = (M + 3) / 4 * 4, ldn = (N + 3) / 4 * 4;
    const int64_t nrb = cdiv(M, kTB), ncb = cdiv(N, kTB), g = cdiv(M, kRPB) + 1;
    const size_t pe = std::max((size_t)nrb * N, (size_t)ncb * M);
    const size_t tr = m < n ? (size_t)ldm * N : 0;   // Aᵀ for the wide case
    return 2 * (size_t)ldm * kPanel + 2 * (size_t)ldn * kPanel + pe +
           2 * (size_t)nrb * kPanel + 2 * (size_t)ncb * kPanel + 2 * (size_t)g + 16 + tr;
}

// bidiagonalize! of a tall A (m ≥ n).
template <typename T>
int bidiag_tall_dev(hipStream_t s, int64_t m, int64_t n, T *A, int64_t lda, T *taul, T *taur,
                    T *ws)
{
    const int64_t ldm = (m + 3) / 4 * 4, ldn = (n + 3) / 4 * 4;
    const int64_t nrb_max = cdiv(m, kTB), ncb_max = cdiv(n, kTB), gmax = cdiv(m, kRPB) + 1;
    T *V = ws, *X = V + ldm * kPanel, *Y = X + ldm * kPanel, *U = Y + ldn * kPanel;
    T *P = U + ldn * kPanel;
    T *Cv1 = P + std::max(nrb_max * n, ncb_max * m), *Cv2 = Cv1 + nrb_max * kPanel;
    T *Cu1 = Cv2 + nrb_max * kPanel, *Cu2 = Cu1 + ncb_max * kPanel;
    T *npart = Cu2 + ncb_max * kPanel, *npart2 = npart + gmax, *scal = npart2 + gmax;
    for (int64_t k = 0; k < n; k += kPanel) {
        const int w = (int)std::min<int64_t>(kPanel, n - k);
        int ncnt = (int)cdiv(m - k, 256);
        hipLaunchKernelGGL((bd_col0_kernel<T>), dim3((unsigned)ncnt), dim3(256), 0, s, m, k,
                           (const T *)A, lda, npart);
        GLA_LAUNCH_CHECK();
        for (int j = 0; j < w; ++j) {
            const int64_t i = k + j;
            const int64_t C = n - i - 1, R = m - i;
            const int64_t nrb = cdiv(R, kTB), ncb = std::max<int64_t>(1, cdiv(C, kTB));
            const unsigned g2 = (unsigned)(C > 0 ? nrb * ncb : 1);
            hipLaunchKernelGGL((bd_gemvt_kernel<T>), dim3(g2), dim3(256), 0, s, m, n, i, j,
                               (const T *)A, lda, (const T *)V, (const T *)X, ldm,
                               (const T *)npart, ncnt, ncb, P, Cv1, Cv2, scal, taul);
            GLA_LAUNCH_CHECK();
            const unsigned g3 = (unsigned)cdiv(std::max(R, C), kRPB);
            hipLaunchKernelGGL((bd_y_kernel<T>), dim3(g3), dim3(256), 0, s, m, n, i, j, A, lda, V,
                               (const T *)X, ldm, Y, (const T *)U, ldn, (const T *)P,
                               C > 0 ? nrb : 0, (const T *)Cv1, (const T *)Cv2, (const T *)scal,
                               npart2);
            GLA_LAUNCH_CHECK();
            if (C == 0) break;   // the last column: no right reflector (src/svd.jl:339)
            const int64_t R1 = m - i - 1;
            const int64_t nrb1 = cdiv(R1, kTB), ncb1 = cdiv(C, kTB);
            hipLaunchKernelGGL((bd_gemvn_kernel<T>), dim3((unsigned)(nrb1 * ncb1)), dim3(256), 0,
                               s, m, n, i, j, (const T *)A, lda, (const T *)Y, (const T *)U, ldn,
                               (const T *)npart2, (int)g3, ncb1, P, Cu1, Cu2, scal, taur);
            GLA_LAUNCH_CHECK();
            const int last = j == w - 1;
            const unsigned g5 = (unsigned)cdiv(std::max(R1, C), kRPB);
            hipLaunchKernelGGL((bd_x_kernel<T>), dim3(g5), dim3(256), 0, s, m, n, i, j, last, A,
                               lda, (const T *)V, X, ldm, (const T *)Y, U, ldn, (const T *)P,
                               ncb1, (const T *)Cu1, (const T *)Cu2, (const T *)scal, npart);
            GLA_LAUNCH_CHECK();
            ncnt = (int)g5;
        }
        const int64_t c1 = k + w;
        if (c1 < n) {   // Ã[c1:, c1:] −= V Yᵀ + X Uᵀ
            GLA_TRY(gemm<T>(s, 'N', 'T', 0, m - c1, n - c1, w, T(-1), V + c1, ldm, Y + c1, ldn,
                            T(1), A + c1 + c1 * lda, lda, 0));
            GLA_TRY(gemm<T>(s, 'N', 'T', 0, m - c1, n - c1, w, T(-1), X + c1, ldm, U + c1, ldn,
                            T(1), A + c1 + c1 * lda, lda, 0));
        }
    }
    return GLA_OK;
}

// bidiagonalize!: tall directly, wide (m < n) as the tall problem of Aᵀ
// (its left reflectors are A's right ones, src/svd.jl:356-375 mirrors
// :332-349), transposed back.
template <typename T>
int bidiag_dev(hipStream_t s, int64_t m, int64_t n, T *A, int64_t lda, T *taul, T *taur, T *ws)
{
    if (m == 0 || n == 0) return GLA_OK;
    if (m >= n) return bidiag_tall_dev<T>(s, m, n, A, lda, taul, taur, ws);
    const int64_t ldt = (n + 3) / 4 * 4;
    T *At = ws;
    T *rest = ws + (size_t)ldt * m;
    hipLaunchKernelGGL((transpose_kernel<T>), dim3((unsigned)cdiv(m, 16), (unsigned)cdiv(n, 16)),
                       dim3(16, 16), 0, s, m, n, (const T *)A, lda, At, ldt);
    GLA_LAUNCH_CHECK();
    GLA_TRY(bidiag_tall_dev<T>(s, n, m, At, ldt, taur, taul, rest));
    hipLaunchKernelGGL((transpose_kernel<T>), dim3((unsigned)cdiv(n, 16), (unsigned)cdiv(m, 16)),
                       dim3(16, 16), 0, s, n, m, (const T *)At, ldt, A, lda);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

template int bidiag_dev<double>(hipStream_t, int64_t, int64_t, double *, int64_t, double *,
                                double *, double *);
template int bidiag_dev<float>(hipStream_t, int64_t, int64_t, float *, int64_t, float *, float *,
                               float *);

}  // namespace gla

// ------------------------------------------------------------------ C ABI --
namespace {

template <typename T>
int symtri_entry(gla_handle_t hh, char uplo, int64_t n, T *A, int64_t lda, T *tau, void *stream)
{
    gla::Handle *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U') { gla::set_error("uplo must be 'L' or 'U'"); return -2; }
    if (n < 0) { gla::set_error("n < 0"); return -3; }
    if (!A && n > 0) { gla::set_error("A is null"); return -4; }
    if (lda < std::max<int64_t>(1, n)) { gla::set_error("lda < max(1,n)"); return -5; }
    if (!tau && n > 1) { gla::set_error("tau is null"); return -6; }
    if (n == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(gla::ensure_workspace(h, gla::symtri_scratch(n) * sizeof(T)));
    return gla::symtri_dev<T>(reinterpret_cast<hipStream_t>(stream), uplo, n, A, lda, tau,
                              reinterpret_cast<T *>(h->ws));
}


template <typename T>
int bidiag_entry(gla_handle_t hh, int64_t m, int64_t n, T *A, int64_t lda, T *taul, T *taur,
                 void *stream)
{
    gla::Handle *h = reinterpret_cast<gla::Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (m < 0) { gla::set_error("m < 0"); return -2; }
    if (n < 0) { gla::set_error("n < 0"); return -3; }
    if (!A && m > 0 && n > 0) { gla::set_error("A is null"); return -4; }
    if (lda < std::max<int64_t>(1, m)) { gla::set_error("lda < max(1,m)"); return -5; }
    const int64_t k = std::min(m, n);
    const int64_t nl = m >= n ? k : k - 1, nr = m >= n ? k - 1 : k;
    if (!taul && nl > 0) { gla::set_error("taul is null"); return -6; }
    if (!taur && nr > 0) { gla::set_error("taur is null"); return -7; }
    if (k == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(gla::ensure_workspace(h, gla::bidiag_scratch(m, n) * sizeof(T)));
    return gla::bidiag_dev<T>(reinterpret_cast<hipStream_t>(stream), m, n, A, lda, taul, taur,
                              reinterpret_cast<T *>(h->ws));
}

}  // namespace

extern "C" {

int gla_symtri_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda, double *tau,
                 void *stream)
{
    return symtri_entry<double>(h, uplo, n, A, lda, tau, stream);
}
int gla_symtri_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda, float *tau,
                 void *stream)
{
    return symtri_entry<float>(h, uplo, n, A, lda, tau, stream);
}

int gla_bidiag_d(gla_handle_t h, int64_t m, int64_t n, double *A, int64_t lda, double *taul,
                 double *taur, void *stream)
{
    return bidiag_entry<double>(h, m, n, A, lda, taul, taur, stream);
}
int gla_bidiag_s(gla_handle_t h, int64_t m, int64_t n, float *A, int64_t lda, float *taul,
                 float *taur, void *stream)
{
    return bidiag_entry<float>(h, m, n, A, lda, taul, taur, stream);
}

}  // extern "C"
