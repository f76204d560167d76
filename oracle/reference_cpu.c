/*
 * The reference's CPU path, for the benchmark's CPU baseline: cholRecursive!
 * (src/cholesky.jl:37-55) and qrBlocked! (src/qr.jl:113-146) on Float64 as
 * GenericLinearAlgebra.jl runs them.  The package's own generic loops run
 * single-threaded, as Julia runs them (the bodies of oracle_generic.h,
 * compiled here with loop vectorisation, as Julia's LLVM compiles them);
 * the steps Julia's LinearAlgebra hands to BLAS/LAPACK run on an OpenBLAS
 * the caller supplies through function pointers (ILP64 CBLAS/LAPACK
 * entry points), with its threads:
 *   rdiv!(A21, LowerTriangular(A11)')      src/cholesky.jl:48   -> dtrsm
 *   dot(view, view)                        src/qr.jl:72         -> ddot
 *   inv!(UnitUpperTriangular(T))           src/qr.jl:75         -> dtrtri
 *   lmul!(V1', M), lmul!(H.T', M)          src/householder.jl:142,146 -> dtrmm
 *   mul!(M, V2', A2, 1, 1), mul!(A2, V2, M, -1, 1)   :144,149   -> dgemm
 * and the generic ones stay loops:
 *   rankUpdate!(Hermitian(A22,:L), A21, -1)  src/juliaBLAS.jl:89-112
 *   qrUnblocked! (reflector!, reflectorApply!)  src/qr.jl:86-111
 *   lmul!(V1, M, -1)                           src/juliaBLAS.jl:154-166
 *   axpy!(1, M, A1) on matrix views            src/householder.jl:154
 *
 * TEST AND BENCHMARK INFRASTRUCTURE ONLY: nothing in the product links it.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define OT double
#define OFN(name) OCAT(name, r)
#define OSQRT sqrt
#define OCOPYSIGN copysign
#include "oracle_generic.h"
#undef OT
#undef OFN
#undef OSQRT
#undef OCOPYSIGN

#define IX(i, j, ld) ((int64_t)(i) + (int64_t)(j) * (int64_t)(ld))

/* CBLAS constants */
enum { kColMajor = 102, kNoTrans = 111, kTrans = 112, kUpper = 121, kLower = 122,
       kNonUnit = 131, kUnit = 132, kLeft = 141, kRight = 142 };

typedef void (*trsm_fn)(int, int, int, int, int, int64_t, int64_t, double, const double *,
                        int64_t, double *, int64_t);
typedef void (*trmm_fn)(int, int, int, int, int, int64_t, int64_t, double, const double *,
                        int64_t, double *, int64_t);
typedef void (*gemm_fn)(int, int, int, int64_t, int64_t, int64_t, double, const double *,
                        int64_t, const double *, int64_t, double, double *, int64_t);
typedef double (*dot_fn)(int64_t, const double *, int64_t, const double *, int64_t);
typedef void (*trtri_fn)(const char *, const char *, const int64_t *, double *, const int64_t *,
                         int64_t *, size_t, size_t);

static struct {
    trsm_fn trsm;
    trmm_fn trmm;
    gemm_fn gemm;
    dot_fn dot;
    trtri_fn trtri;
} blas;

/* Installs the BLAS/LAPACK entry points; 0 on success. */
int orf_set_blas(void *trsm, void *trmm, void *gemm, void *dot, void *trtri)
{
    blas.trsm = (trsm_fn)trsm;
    blas.trmm = (trmm_fn)trmm;
    blas.gemm = (gemm_fn)gemm;
    blas.dot = (dot_fn)dot;
    blas.trtri = (trtri_fn)trtri;
    return (trsm && trmm && gemm && dot && trtri) ? 0 : -1;
}

/* cholRecursive!(A, Val{:L}, cutoff) — src/cholesky.jl:37-55.  Returns 0,
 * or the 1-based index of the pivot whose sqrt would throw DomainError (the
 * factorisation stops there, as the exception does). */
int64_t orf_chol_recursive(int64_t n, double *A, int64_t lda, int64_t cutoff)
{
    if (n == 1) {                                                      /* :39-40 */
        if (A[0] < 0) return 1;
        A[0] = sqrt(A[0]);
        return 0;
    }
    if (n < cutoff) return orc_chol_unblocked_r(n, A, lda);          /* :41-42 */
    const int64_t n2 = n / 2;                                          /* :44 */
    int64_t info = orf_chol_recursive(n2, A, lda, 1);                  /* :46 */
    if (info) return info;
    blas.trsm(kColMajor, kRight, kLower, kTrans, kNonUnit, n - n2, n2, 1.0, A, lda, A + n2,
              lda);                                                    /* :47-48 */
    orc_syrk_l_r(n - n2, n2, -1.0, A + n2, lda, A + IX(n2, n2, lda), lda); /* :50-51 */
    info = orf_chol_recursive(n - n2, A + IX(n2, n2, lda), lda, cutoff);   /* :52 */
    return info ? info + n2 : 0;
}

/* F[Tuple{:QBlocked}].T of the panel QR2 (m × k factors) — src/qr.jl:64-83. */
static void orf_tblock(int64_t m, int64_t k, const double *F, int64_t ldf, const double *tau,
                       double *Tm, int64_t ldt)
{
    for (int64_t j = 0; j < k; ++j)
        for (int64_t i = 0; i < k; ++i) Tm[IX(i, j, ldt)] = 0.0;
    for (int64_t j = 0; j < k; ++j)                                   /* :70-74 */
        for (int64_t i = 0; i < j; ++i) {
            const int64_t len = m - j - 1;
            const double d = len > 0 ? blas.dot(len, F + IX(j + 1, i, ldf), 1,
                                                F + IX(j + 1, j, ldf), 1)
                                     : 0.0;
            Tm[IX(i, j, ldt)] = tau[i] * (F[IX(j, i, ldf)] + d);
        }
    int64_t info = 0, kk = k, ld = ldt;
    blas.trtri("U", "U", &kk, Tm, &ld, &info, 1, 1);                  /* :75 */
    for (int64_t j = 0; j < k; ++j) {                                  /* :76-81 */
        Tm[IX(j, j, ldt)] = tau[j];
        for (int64_t i = 0; i < j; ++i) Tm[IX(i, j, ldt)] *= tau[j];
    }
}

/* lmul!(adjoint(H), A, M) — src/householder.jl:119-157, H = (V, T),
 * b = min(mA, size(V, 2)), M b×nA. */
static void orf_wy_apply_adj(int64_t mA, int64_t nA, const double *V, int64_t ldv, int64_t nV,
                             const double *Tm, int64_t ldt, double *A, int64_t lda, double *M,
                             int64_t ldm)
{
    const int64_t b = mA < nV ? mA : nV;                               /* :128 */
    for (int64_t j = 0; j < nA; ++j)                                   /* :140 */
        memcpy(M + IX(0, j, ldm), A + IX(0, j, lda), sizeof(double) * (size_t)b);
    blas.trmm(kColMajor, kLeft, kLower, kTrans, kUnit, b, nA, 1.0, V, ldv, M, ldm);  /* :142 */
    if (mA > b)                                                        /* :144 */
        blas.gemm(kColMajor, kTrans, kNoTrans, b, nA, mA - b, 1.0, V + b, ldv, A + b, lda, 1.0,
                  M, ldm);
    blas.trmm(kColMajor, kLeft, kUpper, kTrans, kNonUnit, b, nA, 1.0, Tm, ldt, M, ldm); /* :146 */
    if (mA > b)                                                        /* :149 */
        blas.gemm(kColMajor, kNoTrans, kNoTrans, mA - b, nA, b, -1.0, V + b, ldv, M, ldm, 1.0,
                  A + b, lda);
    orc_lmul_unit_lower_r(b, nA, V, ldv, M, ldm, -1.0);                /* :153 */
    for (int64_t j = 0; j < nA; ++j)                                   /* :154 */
        for (int64_t i = 0; i < b; ++i) A[IX(i, j, lda)] += M[IX(i, j, ldm)];
}

/* qrBlocked!(A, bs, τ, work) — src/qr.jl:113-146 (the tail recursion as a
 * loop); work is bs × n with ld ldw ≥ bs. */
void orf_qr_blocked(int64_t m, int64_t n, double *A, int64_t lda, int64_t bs, double *tau,
                    double *work, int64_t ldw)
{
    double *Tm = (double *)malloc(sizeof(double) * (size_t)(bs * bs > 0 ? bs * bs : 1));
    int64_t off = 0;
    for (;;) {
        const int64_t mm = m - off, nn = n - off;
        double *Ab = A + IX(off, off, lda);
        const int64_t w = nn < bs ? nn : bs;                           /* :121 */
        orc_qr_unblocked_r(mm, w, Ab, lda, tau + off);                 /* :123-125 */
        if (nn > bs) {                                                 /* :126-133 */
            const int64_t k = mm < w ? mm : w;
            orf_tblock(mm, k, Ab, lda, tau + off, Tm, k);
            orf_wy_apply_adj(mm, nn - bs, Ab, lda, w, Tm, k, Ab + IX(0, bs, lda), lda, work, ldw);
        }
        if (!(mm > bs && nn > bs)) break;                              /* :136 */
        off += bs;
    }
    free(Tm);
}
