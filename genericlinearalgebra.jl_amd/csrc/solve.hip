// Solves with the factors: A X = B from a Cholesky factor (LAPACK potrs,
// Julia's ldiv! with a Cholesky) and the least-squares solution of A X ≈ B
// from a Householder QR (Julia's ldiv! with a QR: lmul!(Hᴴ, B) of
// src/householder.jl:119-157, then back-substitution with R).  The
// reference defines the factorisations only (src/cholesky.jl:3-55,
// src/qr.jl:86-146); these are the solves a caller applies to their
// results.
//
// Device plan.  The triangular solves run as right-side solves on the
// transposed right-hand sides (rows of length n, one per right-hand side)
// with the Cholesky driver's B <- B L⁻ᵀ (16×16 tile inverses, MFMA GEMMs
// between 128-wide leaves):
//   L y = b           ⇔  yᵀ = bᵀ L⁻ᵀ;
//   Lᵀ x = y          ⇔  xᵀ = yᵀ L⁻¹ = ((yᵀ J) Lr⁻ᵀ) J,  Lr = J Lᵀ J lower;
//   R x = y (R upper) ⇔  xᵀ = yᵀ R⁻ᵀ = ((yᵀ J) Kr⁻ᵀ) J,  Kr = J R J lower,
// J the reversal (Lr⁻ᵀ = J L⁻¹ J).  Qᵀ B is the compact-WY apply of the
// factorisation's own reflectors.
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

inline dim3 g2(int64_t rows, int64_t cols)
{
    return dim3((unsigned)cdiv(rows, 256), (unsigned)std::min<int64_t>(cols, 65535));
}

// Y (nrhs×n, row k = right-hand side k) from B (n×nrhs):
// Y[k][i] = B[rev ? n−1−i : i][k].
template <typename T>
__global__ void to_rows_k(int64_t n, int64_t nrhs, const T *B, int64_t ldb, T *Y, int rev)
{
    GLA_ROW(i);
    GLA_COLS(k, nrhs)
    if (i < n) Y[k + i * nrhs] = B[(rev ? n - 1 - i : i) + k * ldb];
}

// Z = Y J for Y nrhs×n: Z[k][j] = Y[k][n−1−j].
template <typename T>
__global__ void colrev_k(int64_t n, int64_t nrhs, const T *Y, T *Z)
{
    GLA_ROW(k);
    GLA_COLS(j, n)
    if (k < nrhs) Z[k + j * nrhs] = Y[k + (n - 1 - j) * nrhs];
}

// B (n×nrhs) = (Z J)ᵀ: B[i][k] = Z[k][n−1−i].
template <typename T>
__global__ void from_rows_rev_k(int64_t n, int64_t nrhs, const T *Z, T *B, int64_t ldb)
{
    GLA_ROW(i);
    GLA_COLS(k, nrhs)
    if (i < n) B[i + k * ldb] = Z[k + (n - 1 - i) * nrhs];
}

// Lower-triangular reversals (strict upper part not written):
// transpose = 1: D = J Sᵀ J of a lower S;  transpose = 0: D = J S J of an upper S.
template <typename T>
__global__ void rev_lower_k(int64_t n, const T *S, int64_t lds, T *D, int64_t ldd, int transpose)
{
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n && i >= j)
        D[i + j * ldd] = transpose ? S[(n - 1 - j) + (n - 1 - i) * lds]
                                   : S[(n - 1 - i) + (n - 1 - j) * lds];
}

// D (lower) = Sᵀ of an upper S.
template <typename T>
__global__ void upper_to_lower_k(int64_t n, const T *S, int64_t lds, T *D, int64_t ldd)
{
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n && i >= j) D[i + j * ldd] = S[j + i * lds];
}

// *info = 1 + the first index with a zero diagonal entry of R (unchanged if
// none; the caller sets it to n + 1 first).
template <typename T>
__global__ void first_zero_diag_k(int64_t n, const T *R, int64_t ldr, int *info)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && R[i + i * ldr] == T(0)) atomicMin(info, (int)(i + 1));
}

__global__ void set_int_k(int *p, int v) { *p = v; }

// S (rows×cols, ld lds) = 0.
template <typename T>
__global__ void set_zero_k(int64_t rows, int64_t cols, T *S, int64_t lds)
{
    GLA_ROW(i);
    GLA_COLS(j, cols)
    if (i < rows) S[i + j * lds] = T(0);
}

// Y (n×nrhs, column-major) <- D Y, D the diagonal of F.
template <typename T>
__global__ void scale_rows_k(int64_t n, int64_t nrhs, const T *F, int64_t ldf, T *Y, int64_t ldy)
{
    GLA_ROW(i);
    GLA_COLS(k, nrhs)
    if (i < n) Y[i + k * ldy] *= F[i + i * ldf];
}

// Lx (n×n, lower, explicit unit diagonal) = the unit lower factor of LDLᵀ
// data F: L = F's strict lower triangle (upper = 0), or Uᵀ (upper = 1).
template <typename T>
__global__ void unit_lower_k(int64_t n, const T *F, int64_t ldf, T *Lx, int upper)
{
    GLA_ROW(i);
    GLA_COLS(j, n)
    if (i < n && i >= j)
        Lx[i + j * n] = i == j ? T(1) : (upper ? F[j + i * ldf] : F[i + j * ldf]);
}

// Y (nrhs×n, row k = right-hand side k): Y[k][i] /= d_i, d the diagonal of F.
template <typename T>
__global__ void div_rows_k(int64_t n, int64_t nrhs, const T *F, int64_t ldf, T *Y)
{
    GLA_ROW(k);
    GLA_COLS(i, n)
    if (k < nrhs) Y[k + i * nrhs] = Y[k + i * nrhs] / F[i + i * ldf];
}

#undef GLA_COLS
#undef GLA_ROW

// With Z = yᵀ J (nrhs×n): Z <- Z M⁻ᵀ for the lower M, then B[0:n] = (Z J)ᵀ.
template <typename T>
int reversed_tail(hipStream_t s, int64_t n, int64_t nrhs, const T *M, int64_t ldm, T *Minv,
                  T *Z, T *B, int64_t ldb)
{
    GLA_TRY(tile_inverses<T>(s, n, M, ldm, Minv));
    GLA_TRY(trsm_rlt<T>(s, nrhs, n, M, ldm, Minv, Z, nrhs));
    hipLaunchKernelGGL((from_rows_rev_k<T>), g2(n, nrhs), dim3(256), 0, s, n, nrhs, Z, B, ldb);
    GLA_LAUNCH_CHECK();
    return GLA_OK;
}

}  // namespace

// Scratch elements of the Cholesky solve: Uᵀ (uplo 'U'), Lr, two tile-inverse
// sets, two nrhs×n row blocks.
size_t chol_solve_scratch(int64_t n, int64_t nrhs)
{
    return 2 * (size_t)n * n + 2 * tile_inverse_elems(n) + 2 * (size_t)nrhs * n + 256;
}

// B (n×nrhs) <- A⁻¹ B with A = L Lᵀ (uplo 'L', L lower) or A = Uᵀ U
// (uplo 'U', U upper); the other triangle of F is not read.
template <typename T>
int chol_solve_dev(hipStream_t s, char uplo, int64_t n, int64_t nrhs, const T *F, int64_t ldf,
                   T *B, int64_t ldb, T *scratch)
{
    if (n <= 0 || nrhs <= 0) return GLA_OK;
    const size_t nn = (size_t)n * n, ti = tile_inverse_elems(n);
    T *Lx = scratch, *Lr = Lx + nn, *Li = Lr + nn, *Lri = Li + ti, *Y = Lri + ti,
      *Z = Y + (size_t)nrhs * n;
    const T *L = F;
    int64_t ldl = ldf;
    if (uplo == 'U') {   // L = Uᵀ
        hipLaunchKernelGGL((upper_to_lower_k<T>), g2(n, n), dim3(256), 0, s, n, F, ldf, Lx, n);
        GLA_LAUNCH_CHECK();
        L = Lx;
        ldl = n;
    }
    // yᵀ = bᵀ L⁻ᵀ
    hipLaunchKernelGGL((to_rows_k<T>), g2(n, nrhs), dim3(256), 0, s, n, nrhs, B, ldb, Y, 0);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses<T>(s, n, L, ldl, Li));
    GLA_TRY(trsm_rlt<T>(s, nrhs, n, L, ldl, Li, Y, nrhs));
    // xᵀ = ((yᵀ J) Lr⁻ᵀ) J
    hipLaunchKernelGGL((colrev_k<T>), g2(nrhs, n), dim3(256), 0, s, n, nrhs, Y, Z);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((rev_lower_k<T>), g2(n, n), dim3(256), 0, s, n, L, ldl, Lr, n, 1);
    GLA_LAUNCH_CHECK();
    return reversed_tail<T>(s, n, nrhs, Lr, n, Lri, Z, B, ldb);
}

// B (n×nrhs) <- L⁻ᵀ D⁻¹ L⁻¹ B for LDLᵀ factor data F (uplo 'L': L in the
// strict lower triangle, D on the diagonal; 'U': L = Uᵀ) — ldiv!(F::LDLt, X),
// src/ldlt.jl:225-228.  Same scratch as the Cholesky solve.
template <typename T>
int ldlt_solve_dev(hipStream_t s, char uplo, int64_t n, int64_t nrhs, const T *F, int64_t ldf,
                   T *B, int64_t ldb, T *scratch)
{
    if (n <= 0 || nrhs <= 0) return GLA_OK;
    const size_t nn = (size_t)n * n, ti = tile_inverse_elems(n);
    T *Lx = scratch, *Lr = Lx + nn, *Li = Lr + nn, *Lri = Li + ti, *Y = Lri + ti,
      *Z = Y + (size_t)nrhs * n;
    hipLaunchKernelGGL((unit_lower_k<T>), g2(n, n), dim3(256), 0, s, n, F, ldf, Lx,
                       uplo == 'U' ? 1 : 0);
    GLA_LAUNCH_CHECK();
    // yᵀ = bᵀ L⁻ᵀ (ldiv!(L, X)), then ldiv!(D, X)
    hipLaunchKernelGGL((to_rows_k<T>), g2(n, nrhs), dim3(256), 0, s, n, nrhs, B, ldb, Y, 0);
    GLA_LAUNCH_CHECK();
    GLA_TRY(tile_inverses<T>(s, n, Lx, n, Li));
    GLA_TRY(trsm_rlt<T>(s, nrhs, n, Lx, n, Li, Y, nrhs));
    hipLaunchKernelGGL((div_rows_k<T>), g2(nrhs, n), dim3(256), 0, s, n, nrhs, F, ldf, Y);
    GLA_LAUNCH_CHECK();
    // ldiv!(L', X): xᵀ = ((yᵀ J) Lr⁻ᵀ) J
    hipLaunchKernelGGL((colrev_k<T>), g2(nrhs, n), dim3(256), 0, s, n, nrhs, Y, Z);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((rev_lower_k<T>), g2(n, n), dim3(256), 0, s, n, Lx, n, Lr, n, 1);
    GLA_LAUNCH_CHECK();
    return reversed_tail<T>(s, n, nrhs, Lr, n, Lri, Z, B, ldb);
}

// B (n×nrhs) <- L D Lᵀ B — lmul!(F::LDLt, X), src/ldlt.jl:229-232 (and
// F * X, :234): two MFMA products with the explicit unit factor.
// Scratch: n² + n·nrhs elements.
template <typename T>
int ldlt_lmul_dev(hipStream_t s, char uplo, int64_t n, int64_t nrhs, const T *F, int64_t ldf,
                  T *B, int64_t ldb, T *scratch)
{
    if (n <= 0 || nrhs <= 0) return GLA_OK;
    T *Lx = scratch, *Y = Lx + (size_t)n * n;
    hipLaunchKernelGGL((set_zero_k<T>), g2(n, n), dim3(256), 0, s, n, n, Lx, n);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((unit_lower_k<T>), g2(n, n), dim3(256), 0, s, n, F, ldf, Lx,
                       uplo == 'U' ? 1 : 0);
    GLA_LAUNCH_CHECK();
    // Y = Lᵀ B (lmul!(L', X)), then D Y (lmul!(D, X)), then B = L Y (lmul!(L, X))
    GLA_TRY(gemm<T>(s, 'T', 'N', 0, n, nrhs, n, T(1), Lx, n, B, ldb, T(0), Y, n, 0));
    hipLaunchKernelGGL((scale_rows_k<T>), g2(n, nrhs), dim3(256), 0, s, n, nrhs, F, ldf, Y, n);
    GLA_LAUNCH_CHECK();
    return gemm<T>(s, 'N', 'N', 0, n, nrhs, n, T(1), Lx, n, Y, n, T(0), B, ldb, 0);
}

// Scratch elements of the QR solve.
size_t qr_solve_scratch(int64_t m, int64_t n, int64_t nrhs)
{
    return (size_t)n * n + std::max(qr_tblock_scratch(m, n), qr_apply_scratch(m, nrhs, n)) +
           (size_t)n * n + tile_inverse_elems(n) + (size_t)nrhs * n + 256;
}

// Least squares for m ≥ n: B (m×nrhs) <- Qᵀ B, then B[0:n] <- R⁻¹ B[0:n],
// with (F, τ) as stored by qrBlocked!/qrUnblocked!.  info: 1 + the first
// zero diagonal entry of R (device int), n + 1 when R is nonsingular.
template <typename T>
int qr_solve_dev(hipStream_t s, int64_t m, int64_t n, int64_t nrhs, const T *F, int64_t ldf,
                 const T *tau, T *B, int64_t ldb, int *info, T *scratch)
{
    if (n <= 0 || nrhs <= 0) return GLA_OK;
    const size_t nn = (size_t)n * n;
    T *Tm = scratch;
    T *rest = Tm + nn;
    T *Kr = rest + std::max(qr_tblock_scratch(m, n), qr_apply_scratch(m, nrhs, n));
    T *Kri = Kr + nn, *Z = Kri + tile_inverse_elems(n);
    hipLaunchKernelGGL(set_int_k, dim3(1), dim3(1), 0, s, info, (int)(n + 1));
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((first_zero_diag_k<T>), dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, n, F,
                       ldf, info);
    GLA_LAUNCH_CHECK();
    GLA_TRY(qr_tblock_dev<T>(s, m, n, F, ldf, tau, Tm, n, rest));                // :64-83
    GLA_TRY(qr_apply_dev<T>(s, 'T', m, nrhs, n, F, ldf, Tm, n, B, ldb, rest));   // lmul!(Hᴴ, B)
    // R x = y: Z = yᵀ J, Z <- Z Kr⁻ᵀ, x = (Z J)ᵀ
    hipLaunchKernelGGL((to_rows_k<T>), g2(n, nrhs), dim3(256), 0, s, n, nrhs, B, ldb, Z, 1);
    GLA_LAUNCH_CHECK();
    hipLaunchKernelGGL((rev_lower_k<T>), g2(n, n), dim3(256), 0, s, n, F, ldf, Kr, n, 0);
    GLA_LAUNCH_CHECK();
    return reversed_tail<T>(s, n, nrhs, Kr, n, Kri, Z, B, ldb);
}

}  // namespace gla

// ------------------------------------------------------------------ C ABI --
namespace {

using gla::Handle;

template <typename T>
int chol_solve_entry(gla_handle_t hh, char uplo, int64_t n, int64_t nrhs, const T *F, int64_t ldf,
                     T *B, int64_t ldb, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U') { gla::set_error("uplo must be 'L' or 'U'"); return -2; }
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldf < std::max<int64_t>(1, n)) return -6;
    if (ldb < std::max<int64_t>(1, n)) return -8;
    if (n == 0 || nrhs == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    GLA_TRY(gla::ensure_workspace(h, gla::chol_solve_scratch(n, nrhs) * sizeof(T)));
    return gla::chol_solve_dev<T>(reinterpret_cast<hipStream_t>(stream), uplo, n, nrhs, F, ldf, B,
                                  ldb, reinterpret_cast<T *>(h->ws));
}

// mode 0: ldiv!(F::LDLt, B); 1: lmul!(F::LDLt, B).
template <typename T>
int ldlt_apply_entry(gla_handle_t hh, int mode, char uplo, int64_t n, int64_t nrhs, const T *F,
                     int64_t ldf, T *B, int64_t ldb, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (uplo != 'L' && uplo != 'U') { gla::set_error("uplo must be 'L' or 'U'"); return -2; }
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldf < std::max<int64_t>(1, n)) { gla::set_error("ldf < max(1,n)"); return -6; }
    if (ldb < std::max<int64_t>(1, n)) { gla::set_error("ldb < max(1,n)"); return -8; }
    if (n == 0 || nrhs == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t need = mode == 0 ? gla::chol_solve_scratch(n, nrhs)
                                  : (size_t)n * n + (size_t)n * nrhs + 64;
    GLA_TRY(gla::ensure_workspace(h, need * sizeof(T)));
    T *ws = reinterpret_cast<T *>(h->ws);
    if (mode == 0) return gla::ldlt_solve_dev<T>(s, uplo, n, nrhs, F, ldf, B, ldb, ws);
    return gla::ldlt_lmul_dev<T>(s, uplo, n, nrhs, F, ldf, B, ldb, ws);
}

template <typename T>
int qr_solve_entry(gla_handle_t hh, int64_t m, int64_t n, int64_t nrhs, const T *F, int64_t ldf,
                   const T *tau, T *B, int64_t ldb, int64_t *info, void *stream)
{
    Handle *h = reinterpret_cast<Handle *>(hh);
    if (!h) { gla::set_error("null handle"); return -1; }
    if (m < 0) return -2;
    if (n < 0 || n > m) { gla::set_error("the QR solve needs 0 <= n <= m"); return -3; }
    if (nrhs < 0) return -4;
    if (ldf < std::max<int64_t>(1, m)) return -6;
    if (!tau && n > 0) return -7;
    if (ldb < std::max<int64_t>(1, m)) return -9;
    if (info) *info = 0;
    if (n == 0 || nrhs == 0) return GLA_OK;
    GLA_HIP_TRY(hipSetDevice(h->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    GLA_TRY(gla::ensure_workspace(h, gla::qr_solve_scratch(m, n, nrhs) * sizeof(T)));
    GLA_TRY(gla::qr_solve_dev<T>(s, m, n, nrhs, F, ldf, tau, B, ldb, h->dinfo + 8,
                                 reinterpret_cast<T *>(h->ws)));
    if (info) {
        int hv = 0;
        GLA_HIP_TRY(hipMemcpyAsync(&hv, h->dinfo + 8, sizeof(int), hipMemcpyDeviceToHost, s));
        GLA_HIP_TRY(hipStreamSynchronize(s));
        *info = hv > n ? 0 : hv;
    }
    return GLA_OK;
}

}  // namespace

extern "C" {

int gla_chol_solve_d(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const double *F,
                     int64_t ldf, double *B, int64_t ldb, void *stream)
{
    return chol_solve_entry<double>(h, uplo, n, nrhs, F, ldf, B, ldb, stream);
}
int gla_chol_solve_s(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const float *F,
                     int64_t ldf, float *B, int64_t ldb, void *stream)
{
    return chol_solve_entry<float>(h, uplo, n, nrhs, F, ldf, B, ldb, stream);
}
int gla_ldlt_solve_d(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const double *F,
                     int64_t ldf, double *B, int64_t ldb, void *stream)
{
    return ldlt_apply_entry<double>(h, 0, uplo, n, nrhs, F, ldf, B, ldb, stream);
}
int gla_ldlt_solve_s(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const float *F,
                     int64_t ldf, float *B, int64_t ldb, void *stream)
{
    return ldlt_apply_entry<float>(h, 0, uplo, n, nrhs, F, ldf, B, ldb, stream);
}
int gla_ldlt_lmul_d(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const double *F,
                    int64_t ldf, double *B, int64_t ldb, void *stream)
{
    return ldlt_apply_entry<double>(h, 1, uplo, n, nrhs, F, ldf, B, ldb, stream);
}
int gla_ldlt_lmul_s(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const float *F,
                    int64_t ldf, float *B, int64_t ldb, void *stream)
{
    return ldlt_apply_entry<float>(h, 1, uplo, n, nrhs, F, ldf, B, ldb, stream);
}
int gla_qr_solve_d(gla_handle_t h, int64_t m, int64_t n, int64_t nrhs, const double *F,
                   int64_t ldf, const double *tau, double *B, int64_t ldb, int64_t *info,
                   void *stream)
{
    return qr_solve_entry<double>(h, m, n, nrhs, F, ldf, tau, B, ldb, info, stream);
}
int gla_qr_solve_s(gla_handle_t h, int64_t m, int64_t n, int64_t nrhs, const float *F,
                   int64_t ldf, const float *tau, float *B, int64_t ldb, int64_t *info,
                   void *stream)
{
    return qr_solve_entry<float>(h, m, n, nrhs, F, ldf, tau, B, ldb, info, stream);
}

}  // extern "C"
