// Reverse-mode derivatives (adjoints) of the Cholesky and QR factorisations.
// No counterpart in the reference (src/cholesky.jl and src/qr.jl define the
// forward factorisations only).  Formulas:
//
// Cholesky, A = L Lᵀ (Murray 2016, arXiv:1602.07527, eq. 10):
//   P  = Φ(Lᵀ L̄)            Φ: lower triangle with the diagonal halved
//   Ā  = ½ L⁻ᵀ (P + Pᵀ) L⁻¹  (symmetric convention)
//   or tril(2Ā) − diag(Ā)    (lower-only convention: the derivative with
//                             respect to the lower triangle the
//                             factorisation reads)
//   Note ½(P + Pᵀ) = ½ copyltu(Lᵀ L̄), copyltu(M) = tril(M) + tril(M,−1)ᵀ.
// QR, A = Q R, m ≥ n, Q̄ = 0 (Walter & Lehmann 2018; Liao et al. 2019):
//   M = R R̄ᵀ,  Ā = Q [copyltu(M) R⁻ᵀ ; 0]
//
// Device plan.  The triangle–triangle products run on the MFMA GEMM with
// the K range of each tile starting at its first row (kfrom_row), so the
// work is that of the triangular product.  Both right-side solves
// X ← X L⁻¹ use the Cholesky driver's solve B ← B L⁻ᵀ with the reversed
// matrix Lr = J Lᵀ J (J the reversal): X L⁻¹ = ((X J) Lr⁻ᵀ) J.  For QR,
// R⁻ᵀ = (Rᵀ)⁻¹ with Rᵀ lower, reversed as Kr = J R J.  The final Q product
// is the compact-WY apply of the factorisation's own reflectors.
#include <algorithm>

#include "gla_internal.hpp"

namespace gla {

namespace {

#define GLA_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != GLA_OK) return _rc; \
    } while (0)

#define GLA_COLS(j, ncols) for (int64_t j = blockIdx.y; j < (ncols); j += gridDim.y)
#define GLA_ROW(i) const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x
// strided batch along grid z (stride 0: one matrix)
#define GLA_BATCH(P, st) P += (int64_t)blockIdx.z * (st)

inline dim3 g2(int64_t rows, int64_t cols, int64_t batch = 1)
{
    return dim3((unsigned)cdiv(rows, 256), (unsigned)std::min<int64_t>(cols, 65535),
                (unsigned)batch);
}

// D (n×n) <- op(tri(S)), zeros elsewhere.  src_lower: the lower triangle of
// S is taken (else the upper); trans: D = tri(S)ᵀ.
template <typename T>
__global__ void tri_copy(int64_t n, const T *S, int64_t lds, T *D, int64_t ldd, int src_lower,
                         int trans, int64_t sS, int64_t sD)
{
    GLA_BATCH(S, sS);
    GLA_BATCH(D, sD);
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n) {
        const int64_t si = trans ? j : i, sj = trans ? i : j;
        const bool keep = src_lower ? si >= sj : si <= sj;
        D[i + j * ldd] = keep ? S[si + sj * lds] : T(0);
    }
}

// D[i][n-1-j] = scale · M[max(i,j)][min(i,j)]: scale·copyltu(M), columns
// reversed (reads the lower triangle of M only).
template <typename T>
__global__ void sym_lower_rev(int64_t n, const T *M, int64_t ldm, T scale, T *D, int64_t ldd,
                              int64_t sM, int64_t sD)
{
    GLA_BATCH(M, sM);
    GLA_BATCH(D, sD);
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n) {
        const int64_t a = i > j ? i : j, b = i > j ? j : i;
        D[i + (n - 1 - j) * ldd] = scale * M[a + b * ldm];
    }
}

// D = J Sᵀ J: D[i][j] = S[n-1-j][n-1-i], n×n.
template <typename T>
__global__ void rev_transpose(int64_t n, const T *S, int64_t lds, T *D, int64_t ldd, int64_t sS,
                              int64_t sD)
{
    GLA_BATCH(S, sS);
    GLA_BATCH(D, sD);
    GLA_ROW(i);
    GLA_COLS(j, n) if (i < n) D[i + j * ldd] = S[(n - 1 - j) + (n - 1 - i) * lds];
}

// Lower-triangular reversals (the strict upper part is not written):
// transpose = 1: Lr = J Lᵀ J of a lower L, Lr[i][j] = L[n-1-j][n-1-i];
// transpose = 0: Kr = J R J of an upper R,  Kr[i][j] = R[n-1-i][n-1-j].
template <typename T>
__global__ void rev_lower(int64_t n, const T *S, int64_t lds, T *D, int64_t ldd, int transpose,
                          int64_t sS, int64_t sD)
{
    GLA_BATCH(S, sS);
    GLA_BATCH(D, sD);
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n && i >= j)
        D[i + j * ldd] = transpose ? S[(n - 1 - j) + (n - 1 - i) * lds]
                                   : S[(n - 1 - i) + (n - 1 - j) * lds];
}

// Cholesky output from Res2 (Z = J Res2ᵀ = Res2 J in exact arithmetic):
// Z[i][j] = ½(Res2[i][n-1-j] + Res2[j][n-1-i]), exactly symmetric.
// lower_only: 2Z below the diagonal, Z on it, 0 above.
template <typename T>
__global__ void chol_adj_out(int64_t n, const T *R, int64_t ldr, T *Ab, int64_t ldab,
                             int lower_only, int64_t sR, int64_t sAb)
{
    GLA_BATCH(R, sR);
    GLA_BATCH(Ab, sAb);
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n) {
        const T z = T(0.5) * (R[i + (n - 1 - j) * ldr] + R[j + (n - 1 - i) * ldr]);
        T v = z;
        if (lower_only) v = i > j ? T(2) * z : (i == j ? z : T(0));
        Ab[i + j * ldab] = v;
    }
}

// QR: Ab[i][j] = Res[i][n-1-j] for i < n, 0 for n ≤ i < m.
template <typename T>
__global__ void qr_adj_top(int64_t m, int64_t n, const T *R, int64_t ldr, T *Ab, int64_t ldab)
{
    GLA_ROW(i);
    GLA_COLS(j, n) if (i < m) Ab[i + j * ldab] = i < n ? R[i + (n - 1 - j) * ldr] : T(0);
}

#undef GLA_COLS
#undef GLA_ROW
#undef GLA_BATCH

template <typename T>
int launch_tri_copy(hipStream_t s, int64_t n, const T *S, int64_t lds, T *D, int src_lower,
                    int trans, int64_t sS = 0, int64_t batch = 1)
{
    hipLaunchKernelGGL((tri_copy<T>), g2(n, n, batch), dim3(256), 0, s, n, S, lds, D, n,
                       src_lower, trans, sS, n * n);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

}  // namespace

size_t chol_adjoint_scratch(int64_t n) { return 4 * (size_t)n * n + tile_inverse_elems(n) + 64; }

// Ā for one matrix; L, L̄ lower (their strict upper parts are not read).
template <typename T>
int chol_adjoint_dev(hipStream_t s, int64_t n, const T *L, int64_t ldl, const T *Lb,
                     int64_t ldlb, T *Ab, int64_t ldab, int lower_only, T *scratch)
{
    if (n <= 0) return GLA_OK;
    const size_t nn = (size_t)n * n;
    T *Lt = scratch, *Gt = Lt + nn, *W = Gt + nn, *Lr = W + nn, *Linv = Lr + nn;
    GLA_TRY(launch_tri_copy<T>(s, n, L, ldl, Lt, 1, 1));     // Lt = tril(L)ᵀ
    GLA_TRY(launch_tri_copy<T>(s, n, Lb, ldlb, Gt, 1, 1));   // Gt = tril(L̄)ᵀ
    // W (lower) = Lt · Gtᵀ = Lᵀ L̄, k from each tile's first row
    GLA_TRY(gemm<T>(s, 'N', 'T', 1, n, n, n, T(1), Lt, n, Gt, n, T(0), W, n, 0, nullptr, 0, 1));
    // Lt <- ½ copyltu(W) J
    hipLaunchKernelGGL((sym_lower_rev<T>), g2(n, n), dim3(256), 0, s, n, W, n, T(0.5), Lt, n,
                       (int64_t)0, (int64_t)0);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((rev_lower<T>), g2(n, n), dim3(256), 0, s, n, L, ldl, Lr, n, 1, (int64_t)0,
                       (int64_t)0);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses<T>(s, n, Lr, n, Linv));
    GLA_TRY(trsm_rlt<T>(s, n, n, Lr, n, Linv, Lt, n));        // Res1 = Psym J Lr⁻ᵀ
    hipLaunchKernelGGL((rev_transpose<T>), g2(n, n), dim3(256), 0, s, n, Lt, n, W, n, (int64_t)0,
                       (int64_t)0);
    GLA_LAUNCH_CHECK();
    GLA_TRY(trsm_rlt<T>(s, n, n, Lr, n, Linv, W, n));         // Res2 = J Res1ᵀ J Lr⁻ᵀ
    hipLaunchKernelGGL((chol_adj_out<T>), g2(n, n), dim3(256), 0, s, n, W, n, Ab, ldab,
                       lower_only, (int64_t)0, (int64_t)0);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

// Per-matrix scratch of the batched form (n ≤ 128): L, Lᵀ, L̄ᵀ, W, Lr and
// the tile inverses.
size_t chol_adjoint_batched_scratch(int64_t n)
{
    return 5 * (size_t)n * n + tile_inverse_elems(n);
}

// Ā for a strided batch of SPD matrices of order n ≤ 128, recomputing each
// factor L from A: the single-matrix sequence above with every step one
// launch over the whole batch.  info[b] = first failing pivot of matrix b.
template <typename T>
int chol_adjoint_batched_dev(hipStream_t s, int64_t n, int64_t batch, const T *A, int64_t lda,
                             int64_t sA, const T *Lb, int64_t ldlb, int64_t sLb, T *Ab,
                             int64_t ldab, int64_t sAb, int lower_only, int *info, T *scratch)
{
    if (n <= 0 || batch <= 0) return GLA_OK;
    const int64_t nn = n * n, sl = (int64_t)tile_inverse_elems(n);
    T *Lw = scratch, *Lt = Lw + batch * nn, *Gt = Lt + batch * nn, *W = Gt + batch * nn;
    T *Lr = W + batch * nn, *Linv = Lr + batch * nn;
    GLA_TRY(launch_tri_copy<T>(s, n, A, lda, Lw, 1, 0, sA, batch));   // L <- tril(A)
    GLA_TRY(chol_factor_batched<T>(s, (int)n, Lw, n, nn, info, Linv, sl, batch));
    GLA_TRY(launch_tri_copy<T>(s, n, Lw, n, Lt, 1, 1, nn, batch));    // Lt = Lᵀ
    GLA_TRY(launch_tri_copy<T>(s, n, Lb, ldlb, Gt, 1, 1, sLb, batch)); // Gt = tril(L̄)ᵀ
    GLA_TRY(gemm_batched<T>(s, 'N', 'T', 1, n, n, n, T(1), Lt, n, nn, Gt, n, nn, T(0), W, n, nn,
                            (int)batch, 1));                          // W = Lᵀ L̄ (lower)
    hipLaunchKernelGGL((sym_lower_rev<T>), g2(n, n, batch), dim3(256), 0, s, n, W, n, T(0.5), Lt,
                       n, nn, nn);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((rev_lower<T>), g2(n, n, batch), dim3(256), 0, s, n, Lw, n, Lr, n, 1, nn,
                       nn);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses_batched<T>(s, n, Lr, n, nn, Linv, sl, batch));
    GLA_TRY(trsm_rlt_batched<T>(s, n, (int)n, Lr, n, nn, Linv, sl, Lt, n, nn, batch));
    hipLaunchKernelGGL((rev_transpose<T>), g2(n, n, batch), dim3(256), 0, s, n, Lt, n, W, n, nn,
                       nn);
    GLA_LAUNCH_CHECK();
    GLA_TRY(trsm_rlt_batched<T>(s, n, (int)n, Lr, n, nn, Linv, sl, W, n, nn, batch));
    hipLaunchKernelGGL((chol_adj_out<T>), g2(n, n, batch), dim3(256), 0, s, n, W, n, Ab, ldab,
                       lower_only, nn, sAb);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

size_t qr_adjoint_scratch(int64_t m, int64_t n)
{
    return 5 * (size_t)n * n + tile_inverse_elems(n) +
           std::max(qr_tblock_scratch(m, n), qr_apply_scratch(m, n, n)) + 64;
}

// Ā (m×n) for A = QR as stored by qrBlocked!/qrUnblocked! (F, τ), R̄ upper.
template <typename T>
int qr_adjoint_dev(hipStream_t s, int64_t m, int64_t n, const T *F, int64_t ldf, const T *tau,
                   const T *Rb, int64_t ldrb, T *Ab, int64_t ldab, T *scratch)
{
    if (n <= 0 || m <= 0) return GLA_OK;
    const size_t nn = (size_t)n * n;
    T *Rz = scratch, *Gz = Rz + nn, *W = Gz + nn, *Kr = W + nn, *Tm = Kr + nn, *Linv = Tm + nn;
    T *rest = Linv + tile_inverse_elems(n);
    GLA_TRY(launch_tri_copy<T>(s, n, F, ldf, Rz, 0, 0));     // R
    GLA_TRY(launch_tri_copy<T>(s, n, Rb, ldrb, Gz, 0, 0));   // triu(R̄)
    // W (lower) = R R̄ᵀ
    GLA_TRY(gemm<T>(s, 'N', 'T', 1, n, n, n, T(1), Rz, n, Gz, n, T(0), W, n, 0, nullptr, 0, 1));
    hipLaunchKernelGGL((sym_lower_rev<T>), g2(n, n), dim3(256), 0, s, n, W, n, T(1), Gz, n,
                       (int64_t)0, (int64_t)0);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((rev_lower<T>), g2(n, n), dim3(256), 0, s, n, Rz, n, Kr, n, 0, (int64_t)0,
                       (int64_t)0);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses<T>(s, n, Kr, n, Linv));
    GLA_TRY(trsm_rlt<T>(s, n, n, Kr, n, Linv, Gz, n));        // copyltu(M) J Kr⁻ᵀ
    hipLaunchKernelGGL((qr_adj_top<T>), g2(m, n), dim3(256), 0, s, m, n, Gz, n, Ab, ldab);
    GLA_LAUNCH_CHECK();
    GLA_TRY(qr_tblock_dev<T>(s, m, n, F, ldf, tau, Tm, n, rest));
    return qr_apply_dev<T>(s, 'N', m, n, n, F, ldf, Tm, n, Ab, ldab, rest);
}

}  // namespace gla

// ------------------------------------------------------------------ C ABI --
namespace {

using gla::Handle;

// (a kernel rather than a memset, so the call can be captured in a graph)
__global__ void zero_ints(int64_t n, int *p)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <typename T>
int chol_adjoint_entry(gla_handle_t hh, char conv, int64_t n, const T *L, int64_t ldl,
                       const T *Lb, int64_t ldlb, T *Ab, int64_t ldab, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (conv != 'S' && conv != 'L') { gla::set_error("convention must be 'S' or 'L'"); return -2; }
    if (n < 0) return -3;
    if (ldl < std::max<int64_t>(1, n)) return -5;
    if (ldlb < std::max<int64_t>(1, n)) return -7;
    if (ldab < std::max<int64_t>(1, n)) return -9;
    if (n == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(gla::ensure_workspace(h, gla::chol_adjoint_scratch(n) * sizeof(T)));
    return gla::chol_adjoint_dev<T>(reinterpret_cast<hipStream_t>(stream), n, L, ldl, Lb, ldlb,
                                    Ab, ldab, conv == 'L', reinterpret_cast<T *>(h->ws));
}

template <typename T>
int chol_adjoint_batched_entry(gla_handle_t hh, char conv, int64_t n, int64_t batch, const T *A,
                               int64_t lda, int64_t sA, const T *Lb, int64_t ldlb, int64_t sLb,
                               T *Ab, int64_t ldab, int64_t sAb, int32_t *info, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (conv != 'S' && conv != 'L') { gla::set_error("convention must be 'S' or 'L'"); return -2; }
    if (n < 0) return -3;
    if (n > 128) {
        gla::set_error("batched Cholesky adjoint: order %lld exceeds 128", (long long)n);
        return GLA_ERR_UNSUPPORTED;
    }
    if (batch < 0) return -4;
    if (lda < std::max<int64_t>(1, n)) return -6;
    if (ldlb < std::max<int64_t>(1, n)) return -9;
    if (ldab < std::max<int64_t>(1, n)) return -12;
    if (n == 0 || batch == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // chunks of at most 65535 matrices (one grid dimension each)
    const int64_t chunk = std::min<int64_t>(batch, 65535);
    const size_t per = gla::chol_adjoint_batched_scratch(n) * sizeof(T);
    const size_t data = ((size_t)chunk * per + 255) / 256 * 256;
    GLA_TRY(gla::ensure_workspace(h, data + 4 * (size_t)batch + 256));
    int *dinfo = info ? reinterpret_cast<int *>(info)
                      : reinterpret_cast<int *>(reinterpret_cast<char *>(h->ws) + data);
    hipLaunchKernelGGL(zero_ints, dim3((unsigned)gla::cdiv(batch, 256)), dim3(256), 0, s, batch,
                       dinfo);
    GLA_LAUNCH_CHECK();
    for (int64_t b0 = 0; b0 < batch; b0 += chunk) {
        const int64_t nb = std::min<int64_t>(chunk, batch - b0);
        GLA_TRY(gla::chol_adjoint_batched_dev<T>(s, n, nb, A + b0 * sA, lda, sA, Lb + b0 * sLb,
                                                 ldlb, sLb, Ab + b0 * sAb, ldab, sAb, conv == 'L',
                                                 dinfo + b0, reinterpret_cast<T *>(h->ws)));
    }
    return GLA_OK;
}

template <typename T>
int qr_adjoint_entry(gla_handle_t hh, int64_t m, int64_t n, const T *F, int64_t ldf,
                     const T *tau, const T *Rb, int64_t ldrb, T *Ab, int64_t ldab, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0 || n > m) { gla::set_error("the QR adjoint needs 0 <= n <= m"); return -3; }
    if (ldf < std::max<int64_t>(1, m)) return -5;
    if (ldrb < std::max<int64_t>(1, n)) return -8;
    if (ldab < std::max<int64_t>(1, m)) return -10;
    if (n == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(gla::ensure_workspace(h, gla::qr_adjoint_scratch(m, n) * sizeof(T)));
    return gla::qr_adjoint_dev<T>(reinterpret_cast<hipStream_t>(stream), m, n, F, ldf, tau, Rb,
                                  ldrb, Ab, ldab, reinterpret_cast<T *>(h->ws));
}

}  // namespace

extern "C" {

int gla_chol_adjoint_d(gla_handle_t h, char conv, int64_t n, const double *L, int64_t ldl,
                       const double *Lbar, int64_t ldlbar, double *Abar, int64_t ldabar,
                       void *stream)
{
    return chol_adjoint_entry<double>(h, conv, n, L, ldl, Lbar, ldlbar, Abar, ldabar, stream);
}
int gla_chol_adjoint_s(gla_handle_t h, char conv, int64_t n, const float *L, int64_t ldl,
                       const float *Lbar, int64_t ldlbar, float *Abar, int64_t ldabar,
                       void *stream)
{
    return chol_adjoint_entry<float>(h, conv, n, L, ldl, Lbar, ldlbar, Abar, ldabar, stream);
}
int gla_chol_adjoint_batched_d(gla_handle_t h, char conv, int64_t n, int64_t batch,
                               const double *A, int64_t lda, int64_t strideA, const double *Lbar,
                               int64_t ldlbar, int64_t strideLbar, double *Abar, int64_t ldabar,
                               int64_t strideAbar, int32_t *info, void *stream)
{
    return chol_adjoint_batched_entry<double>(h, conv, n, batch, A, lda, strideA, Lbar, ldlbar,
                                              strideLbar, Abar, ldabar, strideAbar, info, stream);
}
int gla_chol_adjoint_batched_s(gla_handle_t h, char conv, int64_t n, int64_t batch,
                               const float *A, int64_t lda, int64_t strideA, const float *Lbar,
                               int64_t ldlbar, int64_t strideLbar, float *Abar, int64_t ldabar,
                               int64_t strideAbar, int32_t *info, void *stream)
{
    return chol_adjoint_batched_entry<float>(h, conv, n, batch, A, lda, strideA, Lbar, ldlbar,
                                             strideLbar, Abar, ldabar, strideAbar, info, stream);
}
int gla_qr_adjoint_d(gla_handle_t h, int64_t m, int64_t n, const double *F, int64_t ldf,
                     const double *tau, const double *Rbar, int64_t ldrbar, double *Abar,
                     int64_t ldabar, void *stream)
{
    return qr_adjoint_entry<double>(h, m, n, F, ldf, tau, Rbar, ldrbar, Abar, ldabar, stream);
}
int gla_qr_adjoint_s(gla_handle_t h, int64_t m, int64_t n, const float *F, int64_t ldf,
                     const float *tau, const float *Rbar, int64_t ldrbar, float *Abar,
                     int64_t ldabar, void *stream)
{
    return qr_adjoint_entry<float>(h, m, n, F, ldf, tau, Rbar, ldrbar, Abar, ldabar, stream);
}

}  // extern "C"
