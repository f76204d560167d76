/*
 * gla_hip.h — C ABI of libgla_hip.so, the MI355X (gfx950) implementation of
 * GenericLinearAlgebra.jl's blocked dense factorisations.
 *
 * Conventions (all entry points):
 *   - Matrices are column-major device pointers with an explicit leading
 *     dimension (the Julia layout): element (i,j), 0-based, is A[i + j*lda].
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Work is enqueued on it; the call returns before the GPU finishes unless
 *     a host `info` pointer is given, in which case the call synchronises the
 *     stream and stores the result there.
 *   - Return value: GLA_OK (0), or a negative code: -k for an invalid k-th
 *     argument (LAPACK style), GLA_ERR_HIP for a HIP runtime error,
 *     GLA_ERR_UNSUPPORTED for a shape this build does not handle,
 *     GLA_ERR_TIMEOUT when a bounded device-side wait gave up (the outputs
 *     of that call are then not valid).
 *     gla_last_error_string() describes the last failure on this thread.
 *   - A handle serves one stream at a time: its workspace, grid-barrier
 *     counter, info words and side stream are shared by every call made
 *     with it, so calls that may overlap on the device (different streams,
 *     different host threads) need a handle each.  The reference's
 *     functions keep no state between calls; one handle per Julia task or
 *     stream gives the same re-entrancy.
 *   - `info` (Cholesky): 0, or the 1-based index of the first pivot whose
 *     square root is taken of a negative number — the point where the
 *     reference's sqrt throws DomainError (src/cholesky.jl:5,40).  A zero or
 *     NaN pivot is not an error, as in the reference (Inf/NaN propagate).
 *   - Suffix _d = Float64, _s = Float32.
 *
 * Every entry point names the reference function it replaces.
 */
#ifndef GLA_HIP_H
#define GLA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GLA_OK 0
#define GLA_ERR_HIP (-1000)
#define GLA_ERR_UNSUPPORTED (-1001)
#define GLA_ERR_NOT_SPD (-1002)
#define GLA_ERR_TIMEOUT (-1003)

typedef struct gla_handle_s *gla_handle_t;

/* Complex elements as Julia stores ComplexF64 / ComplexF32: (re, im). */
typedef struct gla_cdouble_s {
    double re, im;
} gla_cdouble;
typedef struct gla_cfloat_s {
    float re, im;
} gla_cfloat;

/* ------------------------------------------------------------- runtime -- */
/* Library state for one device: workspace, side stream, info flag.  No
 * reference counterpart (the reference allocates per call, e.g.
 * `work = Matrix{T}(undef, blocksize, n)` at src/qr.jl:117). */
int gla_handle_create(gla_handle_t *h);
int gla_handle_destroy(gla_handle_t h);
const char *gla_last_error_string(void);
/* Version string of this build (no reference counterpart). */
const char *gla_version(void);
/* Tuning knobs of a handle (no reference counterpart).  "chol_block": a
 * multiple of 128 runs cholRecursive! as a right-looking blocked
 * factorisation with blocks of that width and depth-1 look-ahead; 0 runs it
 * as recursive halving down to 2048-wide leaves; −1 (default) picks from n
 * (0 below 3072, then 256, 512 from 10240, 1024 from 14336).  All compute
 * the same factor up to rounding.  "chol_psolve": 1 (default) solves each
 * look-ahead panel up to 768 wide in one launch (every 64 rows sweep the
 * 128-wide column blocks), 0 by the recursive rdiv!.  "chol_syrk_late": 1 (default) starts the look-ahead
 * driver's trailing SYRK on the side stream after the next block column's
 * update, 0 right after the panel solve (same factor).  "side_cu_reserve":
 * CUs (0 to CUs/2; a new handle takes CUs/8, 32 of 256) that the look-ahead
 * driver's trailing SYRK leaves to the chain of diagonal factor, panel
 * solve and next-block-column update, which then runs on a stream of the
 * handle's own and joins the caller's at the end; 0 runs the chain on the
 * caller's stream and the SYRK on any CU (same factor bit for bit).  "ldlt_block": step width (multiple
 * of 128 in [128, 4096], default 512) of gla_ldlt_{d,s}: 128 runs 128-wide
 * steps, wider ones that width with depth-1 look-ahead (same factor up to
 * rounding).  "mixed_block": block width (multiple of 128,
 * default 512) of the look-ahead driver of the BF16/FP16 factorisation in
 * gla_chol_solve_mixed_d.  "qr_panel": 1 or 2 (default), the form of the
 * Householder panel kernel.  "qr_inner": 32 or 64 factors each QR block
 * column as sub-panels of that width with compact-WY updates between them;
 * 0 factors it in one panel launch; −1 (default) takes 64 for block columns
 * taller than 12288 rows and 0 otherwise.  "qr_lookahead": 1 (default)
 * factors the next QR panel while the side stream applies the current
 * block reflector to the rest of the trailing matrix; 0 runs the
 * reference's sequential loop.  "qr_hr": block columns at least this
 * many rows tall (−1, default: any height for block columns ≥ 64 wide,
 * 2048 rows for narrower ones; 0: none) are factored by two
 * CholeskyQR passes and a reconstruction of the Householder vectors from
 * the orthonormal factor (the LU of S − Q₁ with the sign choice that
 * reflector! makes), checked on the device and falling back to the
 * column-by-column panel when the block column is too ill-conditioned for
 * it.  V, τ and R are the same Householder sequence in every setting.  "gemm_dl" (process-wide): 2 (default) or
 * 1/3 route FP64 NT products on 128×128 tiles (SYRK, rankUpdate!, the
 * trailing updates) to a direct-to-LDS kernel with a 4/3/3-deep ring of
 * 8/8/16-wide K slabs, 0 to the register-staged kernel; "gemm_lds1" (process-wide): 1 (default) stages
 * FP64 128×128 tiles through one LDS buffer, 0 through two.  Same results
 * bit for bit in every setting.  "spin_limit": polls (default 2^25, about a
 * second) before a grid-wide wait of the QR panel gives up and the call
 * returns GLA_ERR_TIMEOUT.  "qr_wait_test": 1 makes every such wait target
 * an arrival count that never comes (a test of the timeout report; the
 * kernels still finish, after spin_limit polls per wait). */
int gla_set_tuning(gla_handle_t h, const char *key, int64_t value);
/* Reads a tuning knob's current value, by the same keys.  No counterpart
 * in the reference. */
int gla_get_tuning(gla_handle_t h, const char *key, int64_t *value);

/* ------------------------------------------------------------ Cholesky -- */
/* cholUnblocked!(A, Val{:L}) — src/cholesky.jl:3-15.
 * uplo 'L' (the reference) or 'U' (extension, U = Lᵀ stored in the upper
 * triangle).  The opposite triangle keeps its input values, as in the
 * reference. */
int gla_chol_unblocked_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda,
                         int64_t *info, void *stream);
int gla_chol_unblocked_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda,
                         int64_t *info, void *stream);

/* cholBlocked!(A, Val{:L}, blocksize) — src/cholesky.jl:17-35. */
int gla_chol_blocked_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream);
int gla_chol_blocked_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream);

/* cholRecursive!(A, Val{:L}, cutoff = 1) — src/cholesky.jl:37-55. */
int gla_chol_recursive_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream);
int gla_chol_recursive_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream);

/* The three Cholesky variants for complex Hermitian A (_z ComplexF64,
 * _c ComplexF32) — src/cholesky.jl:3-55 with the rank updates on her! /
 * herk! (src/juliaBLAS.jl:41-48,78-86): L lower with a real diagonal, the
 * opposite triangle untouched ('U': U = Lᴴ in the upper triangle).  The
 * pivot's square root is the complex one (src/cholesky.jl:5,40), so a
 * pivot with a negative real part is not an error there: *info reports
 * the first such pivot (1-based) and the values after it follow the
 * reference's arithmetic (an imaginary root, then division by its zero
 * real part).  The device steps are 64 wide (cholBlocked!'s blocksize
 * when smaller); the factor does not depend on the step width up to
 * rounding. */
int gla_chol_unblocked_z(gla_handle_t h, char uplo, int64_t n, gla_cdouble *A, int64_t lda,
                         int64_t *info, void *stream);
int gla_chol_unblocked_c(gla_handle_t h, char uplo, int64_t n, gla_cfloat *A, int64_t lda,
                         int64_t *info, void *stream);
int gla_chol_blocked_z(gla_handle_t h, char uplo, int64_t n, gla_cdouble *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream);
int gla_chol_blocked_c(gla_handle_t h, char uplo, int64_t n, gla_cfloat *A, int64_t lda,
                       int64_t blocksize, int64_t *info, void *stream);
int gla_chol_recursive_z(gla_handle_t h, char uplo, int64_t n, gla_cdouble *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream);
int gla_chol_recursive_c(gla_handle_t h, char uplo, int64_t n, gla_cfloat *A, int64_t lda,
                         int64_t cutoff, int64_t *info, void *stream);

/* No counterpart in the reference: cholRecursive!(A, Val{:L}, cutoff) in
 * Float64 for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * handle: the call allocates nothing and runs every launch on `stream`.
 * Its scratch (panel tile inverses and the info word) is one fixed device
 * array per GPU (4.2 MB, held until the process ends), so n is limited to
 * 32768 (larger n: GLA_ERR_UNSUPPORTED, nothing launched).  Calls that
 * overlap on one GPU (other streams, other host threads) share that array
 * and race; serialise them.  Same factor as gla_chol_recursive_d with
 * "chol_block" 0 (bit for bit) and to rounding with its default driver.
 * `info` (host, may be NULL) as for gla_chol_recursive_d; with it the call
 * synchronises the stream (leave it NULL under graph capture). */
int gla_chol_recursive_noalloc_d(char uplo, int64_t n, double *A, int64_t lda, int64_t cutoff,
                                 int64_t *info, void *stream);

/* rankUpdate!(C::Hermitian, A::StridedVecOrMat, α::Real) —
 * src/juliaBLAS.jl:89-112 (and the rank-1 form :53-63 with k = 1).
 * C n×n, only the `uplo` triangle is updated: C += α A Aᵀ, A n×k. */
int gla_rank_update_d(gla_handle_t h, char uplo, int64_t n, int64_t k, double alpha,
                      const double *A, int64_t lda, double *C, int64_t ldc, void *stream);
int gla_rank_update_s(gla_handle_t h, char uplo, int64_t n, int64_t k, float alpha,
                      const float *A, int64_t lda, float *C, int64_t ldc, void *stream);

/* rankUpdate!(C::Hermitian, A, α) for complex C — the herk! the reference
 * dispatches to (src/juliaBLAS.jl:78-86): the uplo triangle of
 * C += α A Aᴴ (α real), the diagonal left real. */
int gla_rank_update_z(gla_handle_t h, char uplo, int64_t n, int64_t k, double alpha,
                      const gla_cdouble *A, int64_t lda, gla_cdouble *C, int64_t ldc, void *stream);
int gla_rank_update_c(gla_handle_t h, char uplo, int64_t n, int64_t k, float alpha,
                      const gla_cfloat *A, int64_t lda, gla_cfloat *C, int64_t ldc, void *stream);

/* rdiv!(B, LowerTriangular(L)') as called at src/cholesky.jl:28,48:
 * B (m×n) <- B L⁻ᵀ, L n×n lower (non-unit). */
int gla_trsm_rlt_d(gla_handle_t h, int64_t m, int64_t n, const double *L, int64_t ldl,
                   double *B, int64_t ldb, void *stream);
int gla_trsm_rlt_s(gla_handle_t h, int64_t m, int64_t n, const float *L, int64_t ldl,
                   float *B, int64_t ldb, void *stream);

/* ------------------------------------------------------------------- QR -- */
/* qrUnblocked!(A, τ) — src/qr.jl:86-111.  A m×n is overwritten with R
 * (upper) and the Householder vectors (strict lower, unit diagonal
 * implicit); tau has min(m,n) entries.  Reflector convention of the Julia
 * stdlib reflector!: β = −copysign(‖x‖, x₁), τ = (x₁−β)/(−β); a column
 * [a,0,…,0] with a ≠ 0 gives τ = 2. */
int gla_qr_unblocked_d(gla_handle_t h, int64_t m, int64_t n, double *A, int64_t lda,
                       double *tau, void *stream);
int gla_qr_unblocked_s(gla_handle_t h, int64_t m, int64_t n, float *A, int64_t lda,
                       float *tau, void *stream);

/* qrBlocked!(A, blocksize = 12, τ, work) — src/qr.jl:113-146.
 * Block columns are min(blocksize, 128) wide; the factors do not depend on
 * the block width (the same Householder sequence, up to rounding).
 * `work` (ld ldwork ≥ min(m, blocksize), at least n − blocksize columns)
 * may be NULL.  When given it is left as the reference leaves it: each step
 * with trailing columns writes its lmul! workspace M = −V₁ Tᵀ (Vᵀ A₂)
 * (src/householder.jl:142-153) to work[0:k, 0:n−j−w], later steps
 * overwriting the top-left; for blocksize > 128 that is the content
 * qrBlocked!(A, 128, τ, work) leaves.
 * Both QR entry points synchronise the stream (to read the panel's
 * wait-timeout flag; GLA_ERR_TIMEOUT if it was raised).  A block column of
 * height m is factored by a register-resident panel whose workgroups must
 * all be resident: narrower sub-panels (64, 32 wide) take block columns up
 * to 512·CUs rows (131072 on MI355X); taller ones return
 * GLA_ERR_UNSUPPORTED. */
int gla_qr_blocked_d(gla_handle_t h, int64_t m, int64_t n, double *A, int64_t lda,
                     int64_t blocksize, double *tau, double *work, int64_t ldwork,
                     void *stream);
int gla_qr_blocked_s(gla_handle_t h, int64_t m, int64_t n, float *A, int64_t lda,
                     int64_t blocksize, float *tau, float *work, int64_t ldwork,
                     void *stream);

/* qrUnblocked!(A, τ) / qrBlocked!(A, blocksize, τ) — src/qr.jl:86-146 —
 * for complex A (_z ComplexF64, _c ComplexF32) with the stdlib's complex
 * reflector! (β = −copysign(‖x‖, real(x₁)) real, τ complex) and Hᴴ =
 * I − τ̄ v vᴴ applied from the left, as qrUnblocked! does.  Block columns
 * are min(blocksize, 32) wide (qrUnblocked!: 32); every width gives
 * qrUnblocked!'s factors, up to rounding.  The reference's own complex
 * qrBlocked! does not: its compact-WY T (src/qr.jl:72) takes F[j,i]
 * where (VᴴV)[i,j] needs conj(F[j,i]), so its trailing updates are not
 * H₁⋯H_k; this build forms the T that is (LAPACK zlarft's).  The
 * workspace is internal (no `work` argument). */
int gla_qr_unblocked_z(gla_handle_t h, int64_t m, int64_t n, gla_cdouble *A, int64_t lda,
                       gla_cdouble *tau, void *stream);
int gla_qr_unblocked_c(gla_handle_t h, int64_t m, int64_t n, gla_cfloat *A, int64_t lda,
                       gla_cfloat *tau, void *stream);
int gla_qr_blocked_z(gla_handle_t h, int64_t m, int64_t n, gla_cdouble *A, int64_t lda,
                     int64_t blocksize, gla_cdouble *tau, void *stream);
int gla_qr_blocked_c(gla_handle_t h, int64_t m, int64_t n, gla_cfloat *A, int64_t lda,
                     int64_t blocksize, gla_cfloat *tau, void *stream);
/* getindex(::QR2, Tuple{:QBlocked}) (src/qr.jl:64-83) for complex factors:
 * T (k×k upper, k = min(m,n)) with I − V T Vᴴ = H₁⋯H_k (see above). */
int gla_qr_tblock_z(gla_handle_t h, int64_t m, int64_t n, const gla_cdouble *F, int64_t ldf,
                    const gla_cdouble *tau, gla_cdouble *T, int64_t ldt, void *stream);
int gla_qr_tblock_c(gla_handle_t h, int64_t m, int64_t n, const gla_cfloat *F, int64_t ldf,
                    const gla_cfloat *tau, gla_cfloat *T, int64_t ldt, void *stream);
/* lmul!(H, A) (trans 'N') and lmul!(H', A) (trans 'C') for a complex
 * HouseholderBlock — src/householder.jl:82-115,119-157: A <- op(I − V T Vᴴ) A,
 * V mA×nV (unit lower, strict-lower storage); workspace internal. */
int gla_wy_apply_z(gla_handle_t h, char trans, int64_t mA, int64_t nA, const gla_cdouble *V,
                   int64_t ldv, int64_t nV, const gla_cdouble *T, int64_t ldt, gla_cdouble *A,
                   int64_t lda, void *stream);
int gla_wy_apply_c(gla_handle_t h, char trans, int64_t mA, int64_t nA, const gla_cfloat *V,
                   int64_t ldv, int64_t nV, const gla_cfloat *T, int64_t ldt, gla_cfloat *A,
                   int64_t lda, void *stream);

/* getindex(::QR2, Tuple{:QBlocked}) — src/qr.jl:64-83: the compact-WY
 * T = (I + D striu(VᵀV))⁻¹ D (k×k upper, k = min(m,n)) of QR2(F, τ). */
int gla_qr_tblock_d(gla_handle_t h, int64_t m, int64_t n, const double *F, int64_t ldf,
                    const double *tau, double *T, int64_t ldt, void *stream);
int gla_qr_tblock_s(gla_handle_t h, int64_t m, int64_t n, const float *F, int64_t ldf,
                    const float *tau, float *T, int64_t ldt, void *stream);

/* lmul!(H::HouseholderBlock, A, M) (trans 'N') and
 * lmul!(adjoint(H), A, M) (trans 'T') — src/householder.jl:82-115,119-157.
 * H = I − V T Vᵀ with V mA×nV (unit lower, strict-lower storage), T b×b,
 * b = min(mA, nV); A mA×nA; M b×nA workspace (may be NULL). */
int gla_wy_apply_d(gla_handle_t h, char trans, int64_t mA, int64_t nA, const double *V,
                   int64_t ldv, int64_t nV, const double *T, int64_t ldt, double *A,
                   int64_t lda, double *M, int64_t ldm, void *stream);
int gla_wy_apply_s(gla_handle_t h, char trans, int64_t mA, int64_t nA, const float *V,
                   int64_t ldv, int64_t nV, const float *T, int64_t ldt, float *A,
                   int64_t lda, float *M, int64_t ldm, void *stream);

/* lmul!(A::UnitLowerTriangular, B, α) — src/juliaBLAS.jl:154-166:
 * B (m×n) <- α L_unit B, L m×m. */
int gla_lmul_unit_lower_d(gla_handle_t h, int64_t m, int64_t n, const double *L,
                          int64_t ldl, double *B, int64_t ldb, double alpha, void *stream);
int gla_lmul_unit_lower_s(gla_handle_t h, int64_t m, int64_t n, const float *L,
                          int64_t ldl, float *B, int64_t ldb, float alpha, void *stream);

/* ----------------------------------------------------------------- LDLᵀ -- */
/* ldlt!(Hermitian(A, uplo), blocksize) — src/ldlt.jl:150-157, the blocked
 * factorisations without pivoting _ldlt_lower_blocked! (:80-104, A = L D Lᵀ:
 * unit L below the diagonal, D on it) and _ldlt_upper_blocked! (:121-148,
 * A = Uᵀ D U: unit U above the diagonal).  Only the `uplo` triangle is read
 * and written; the other keeps its input.  `blocksize` (≥ 1; the
 * reference's default is 128 ÷ sizeof(T)) does not change the factors, up
 * to rounding: the device steps are "ldlt_block" wide (gla_set_tuning,
 * default 512 with look-ahead; 128-wide diagonal blocks).  A zero pivot
 * divides by zero and Inf/NaN propagate, as in the reference. */
int gla_ldlt_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda, int64_t blocksize,
               void *stream);
int gla_ldlt_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda, int64_t blocksize,
               void *stream);
/* ldiv!(F::LDLt, X) — src/ldlt.jl:225-228: B (n×nrhs) <- L⁻ᵀ D⁻¹ L⁻¹ B
 * (uplo 'L') or U⁻¹ D⁻¹ U⁻ᵀ B (uplo 'U') for factor data F from
 * gla_ldlt_{d,s}. */
int gla_ldlt_solve_d(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const double *F,
                     int64_t ldf, double *B, int64_t ldb, void *stream);
int gla_ldlt_solve_s(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const float *F,
                     int64_t ldf, float *B, int64_t ldb, void *stream);
/* lmul!(F::LDLt, X) — src/ldlt.jl:229-232 (and F * X, :234): B <- L D Lᵀ B
 * (uplo 'L') or Uᵀ D U B (uplo 'U'). */
int gla_ldlt_lmul_d(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const double *F,
                    int64_t ldf, double *B, int64_t ldb, void *stream);
int gla_ldlt_lmul_s(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const float *F,
                    int64_t ldf, float *B, int64_t ldb, void *stream);

/* ------------------------------------------------ two-sided reductions -- */
/* symtri!(A::Hermitian/Symmetric) — src/eigenSelfAdjoint.jl:446-564:
 * symtriLower! (uplo 'L') or symtriUpper! (uplo 'U') of a real symmetric
 * A (n×n; only the uplo triangle is read and written).  On return the
 * diagonal holds the tridiagonal's diagonal, the sub- ('L') or super-
 * ('U') diagonal its off-diagonal, and the rest of the triangle the
 * Householder vectors (leading 1 implicit) in the reference's packing;
 * tau (device, n − 1 entries) their τ, the last one 0.  Reflector
 * convention of the Julia stdlib reflector! (as in gla_qr_unblocked_d). */
int gla_symtri_d(gla_handle_t h, char uplo, int64_t n, double *A, int64_t lda, double *tau,
                 void *stream);
int gla_symtri_s(gla_handle_t h, char uplo, int64_t n, float *A, int64_t lda, float *tau,
                 void *stream);

/* bidiagonalize!(A) — src/svd.jl:328-381, real A (m×n).  m ≥ n: upper
 * bidiagonal (diagonal and superdiagonal of A), left reflectors below the
 * diagonal (taul: n entries), right reflectors right of the superdiagonal
 * (taur: n − 1 entries); m < n: lower bidiagonal (diagonal and
 * subdiagonal), right reflectors right of the diagonal (taur: m entries),
 * left reflectors below the subdiagonal (taul: m − 1 entries).  The
 * reference's packing and reflector! convention; taul/taur device vectors
 * (may be NULL when of length 0). */
int gla_bidiag_d(gla_handle_t h, int64_t m, int64_t n, double *A, int64_t lda, double *taul,
                 double *taur, void *stream);
int gla_bidiag_s(gla_handle_t h, int64_t m, int64_t n, float *A, int64_t lda, float *taul,
                 float *taur, void *stream);

/* ------------------------------------------- batched uneven Cholesky -- */
/* No counterpart in the reference (one matrix per call there): factor and
 * solve many SPD matrices of uneven size n ≤ 256 in one persistent launch,
 * each with cholUnblocked!'s result (src/cholesky.jl:3-15) followed by the
 * solve L Lᵀ x = b.  A plan (host-side sizes, sorted into size bins: one
 * workgroup per matrix above `wave_max`, one wave per matrix at or below
 * it, `grid_mult` workgroups per CU) is built once; the batched call then
 * allocates nothing and synchronises nothing, so it can be captured in a
 * hipGraph.  Arrays A, lda, b, info are device arrays of length `batch`
 * (pointers to each matrix / right-hand side, its leading dimension, its
 * info = first negative pivot or 0).  A is overwritten with L (lower), b
 * with x. */
typedef struct gla_batch_plan_s *gla_batch_plan_t;
int gla_batch_plan_create(gla_handle_t h, int64_t batch, const int32_t *n_host, int32_t wave_max,
                          int32_t grid_mult, gla_batch_plan_t *plan);
int gla_batch_plan_destroy(gla_batch_plan_t plan);
int gla_chol_solve_batched_d(gla_handle_t h, gla_batch_plan_t plan, double *const *A,
                             const int32_t *lda, double *const *b, int32_t *info, void *stream);
int gla_chol_solve_batched_s(gla_handle_t h, gla_batch_plan_t plan, float *const *A,
                             const int32_t *lda, float *const *b, int32_t *info, void *stream);

/* ------------------------------- mixed precision + iterative refinement -- */
/* Solves A X = B (A n×n SPD, FP64, lower triangle read; B, X n×nrhs FP64).
 * No counterpart in the reference, whose Cholesky (src/cholesky.jl:3-55)
 * factors in the element type of A only; this extends cholRecursive! with
 * a lower-precision factorisation and FP64 refinement.  `prec` selects the
 * factorisation: 'd' FP64 (the plain path), 's' FP32, 'B' FP32 storage with
 * FP64 diagonal blocks and the trailing updates on BF16 MFMA (FP32
 * accumulation) from each panel split into three BF16 terms, 'H' the same
 * with two FP16 terms of the panel scaled by a power of two, 'b' / 'h' one
 * BF16 / FP16 term (faster; for κ up to about 1e3 / 1e4), 'a' FP32 first,
 * then FP64 if FP32 does not converge (a BF16/FP16 factorisation that
 * breaks down stops within four 512-column blocks of the failing pivot).
 * X is
 * refined in FP64 (r = B − A X with the symmetric A, correction with the
 * factor) until the normwise backward error
 * max_k ‖r_k‖₂ / (‖A‖_F ‖x_k‖₂ + ‖b_k‖₂) ≤ tol; a precision stops when η
 * has not halved over two corrections or after max_iter corrections.
 * *iters = corrections of the last precision tried, *berr = its final
 * backward error (> tol if none converged), *info = first non-positive
 * pivot (1-based) if that factorisation broke down, *prec_used = the
 * precision of the returned X.  A and B are not modified.  Synchronises
 * the stream. */
int gla_chol_solve_mixed_d(gla_handle_t h, char prec, int64_t n, const double *A, int64_t lda,
                           int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx,
                           double tol, int32_t max_iter, int32_t *iters, double *berr,
                           int64_t *info, char *prec_used, void *stream);
/* Low-precision MFMA product used by the mixed factorisation:
 * C (m×n, FP32) += alpha · A Bᵀ with A (m×k), B (n×k) FP32 in memory, each
 * rounded once to `tin` ('b' BF16, 'h' FP16) or split into terms ('B' three
 * BF16 terms, 'H' two FP16 terms after a common power-of-two scaling) in
 * packed copies in the handle's workspace, FP32 accumulation.  Mirrors
 * rankUpdate!'s role (src/juliaBLAS.jl:89). */
int gla_gemm_lowp_nt(gla_handle_t h, char tin, int64_t m, int64_t n, int64_t k, float alpha,
                     const float *A, int64_t lda, const float *B, int64_t ldb, float *C,
                     int64_t ldc, void *stream);

/* --------------------------------------------------------------- solves -- */
/* No counterpart in the reference (it defines the factorisations only,
 * src/cholesky.jl:3-55, src/qr.jl:86-146): the solves a caller applies to
 * their results, Julia's ldiv! with a Cholesky or a QR.
 *
 * Cholesky (LAPACK potrs): B (n×nrhs) <- A⁻¹ B for A = L Lᵀ (uplo 'L', F
 * holds L in its lower triangle — cholRecursive!'s output) or A = Uᵀ U
 * (uplo 'U', U in the upper triangle); the other triangle is not read. */
int gla_chol_solve_d(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const double *F,
                     int64_t ldf, double *B, int64_t ldb, void *stream);
int gla_chol_solve_s(gla_handle_t h, char uplo, int64_t n, int64_t nrhs, const float *F,
                     int64_t ldf, float *B, int64_t ldb, void *stream);
/* QR least squares (m ≥ n), (F, τ) as stored by qrBlocked!/qrUnblocked!:
 * B (m×nrhs) <- Qᵀ B (lmul!(Hᴴ, B), src/householder.jl:119-157), then the
 * first n rows <- R⁻¹ of themselves, the minimiser of ‖A X − B‖_F; rows
 * n..m−1 keep the residual's components.  *info (host, may be NULL): 0, or
 * the 1-based index of the first zero diagonal entry of R (X then holds
 * Inf/NaN).  With info given the call synchronises the stream. */
int gla_qr_solve_d(gla_handle_t h, int64_t m, int64_t n, int64_t nrhs, const double *F,
                   int64_t ldf, const double *tau, double *B, int64_t ldb, int64_t *info,
                   void *stream);
int gla_qr_solve_s(gla_handle_t h, int64_t m, int64_t n, int64_t nrhs, const float *F,
                   int64_t ldf, const float *tau, float *B, int64_t ldb, int64_t *info,
                   void *stream);

/* ------------------------------------------------------------ adjoints -- */
/* No counterpart in the reference (forward factorisations only,
 * src/cholesky.jl, src/qr.jl): reverse-mode derivatives for an rrule.
 *
 * Cholesky: given L (the factor, lower; strict upper not read) and L̄
 * (lower; strict upper not read), Ā (n×n) = ½ L⁻ᵀ (P + Pᵀ) L⁻¹ with
 * P = Φ(Lᵀ L̄), Φ = lower triangle with halved diagonal.  conv 'S': that
 * symmetric matrix; conv 'L': the derivative with respect to the lower
 * triangle the factorisation reads — 2Ā below the diagonal, Ā on it, 0
 * above. */
int gla_chol_adjoint_d(gla_handle_t h, char conv, int64_t n, const double *L, int64_t ldl,
                       const double *Lbar, int64_t ldlbar, double *Abar, int64_t ldabar,
                       void *stream);
int gla_chol_adjoint_s(gla_handle_t h, char conv, int64_t n, const float *L, int64_t ldl,
                       const float *Lbar, int64_t ldlbar, float *Abar, int64_t ldabar,
                       void *stream);
/* No counterpart in the reference: the same for a strided batch of `batch`
 * SPD matrices of order n ≤ 128, recomputing each factor (cholUnblocked!,
 * src/cholesky.jl:3-15): matrix b is A + b·strideA (lower triangle read),
 * its cotangent Lbar + b·strideLbar, its Ā written to Abar + b·strideAbar.
 * info (device, `batch` int32, or NULL): 0, or the first non-positive pivot
 * of that matrix (its Ā is then not meaningful).  Every step of the
 * single-matrix sequence is one launch over the batch. */
int gla_chol_adjoint_batched_d(gla_handle_t h, char conv, int64_t n, int64_t batch,
                               const double *A, int64_t lda, int64_t strideA, const double *Lbar,
                               int64_t ldlbar, int64_t strideLbar, double *Abar, int64_t ldabar,
                               int64_t strideAbar, int32_t *info, void *stream);
int gla_chol_adjoint_batched_s(gla_handle_t h, char conv, int64_t n, int64_t batch,
                               const float *A, int64_t lda, int64_t strideA, const float *Lbar,
                               int64_t ldlbar, int64_t strideLbar, float *Abar, int64_t ldabar,
                               int64_t strideAbar, int32_t *info, void *stream);
/* No counterpart in the reference: QR (m ≥ n), as stored by
 * qrBlocked!/qrUnblocked! (src/qr.jl:86-146; F: R on and above the
 * diagonal, reflectors below; τ), with Q̄ = 0 and R̄ (n×n, upper; strict
 * lower not read): Ā (m×n) = Q [copyltu(R R̄ᵀ) R⁻ᵀ ; 0], copyltu(M) =
 * tril(M) + tril(M,−1)ᵀ, Q the thin factor of the reflectors. */
int gla_qr_adjoint_d(gla_handle_t h, int64_t m, int64_t n, const double *F, int64_t ldf,
                     const double *tau, const double *Rbar, int64_t ldrbar, double *Abar,
                     int64_t ldabar, void *stream);
int gla_qr_adjoint_s(gla_handle_t h, int64_t m, int64_t n, const float *F, int64_t ldf,
                     const float *tau, const float *Rbar, int64_t ldrbar, float *Abar,
                     int64_t ldabar, void *stream);

/* ------------------------------------------------------ GEMM primitive -- */
/* The MFMA GEMM the factorisations are built on, exported for the
 * juliaBLAS-style callers (`mul!(C, A, B, α, β)` at src/householder.jl:108,144,
 * 149): C = α op(A) op(B) + β C, (transa,transb) in {NN, NT, TN}. */
int gla_gemm_d(gla_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
               double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
               double beta, double *C, int64_t ldc, void *stream);
int gla_gemm_s(gla_handle_t h, char transa, char transb, int64_t m, int64_t n, int64_t k,
               float alpha, const float *A, int64_t lda, const float *B, int64_t ldb,
               float beta, float *C, int64_t ldc, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GLA_HIP_H */
