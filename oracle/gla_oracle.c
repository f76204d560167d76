/*
 * CPU oracle for the blocked Cholesky / Householder-QR path of
 * GenericLinearAlgebra.jl (src/cholesky.jl, src/qr.jl, src/householder.jl,
 * src/juliaBLAS.jl).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the test suite,
 * the smoke check and the benchmark's cpu_baseline leg load it; the HIP
 * library never links or calls it.
 *
 * It restates the reference's algorithms in plain C, loop for loop, for
 * float (suffix _s), double (_d) and __float128 (_q; the stand-in for the
 * reference's BigFloat runs).  Julia-stdlib steps the reference calls
 * (reflector!, left reflectorApply!, rdiv!, inv!(UnitUpperTriangular)) are
 * restated from their published algorithms; each function names the lines
 * it follows.
 *
 * Pinning: every entry point is checked in tests/test_oracle.py against
 * LAPACK (scipy dpotrf/spotrf/dgeqrf), closed-form answers (min(i,j),
 * symmetric Pascal, [3;4] reflector, [a;0] reflector with τ=2), explicit
 * products of elementary reflectors, mpmath at 256 bits and brute-force
 * pivots for the DomainError index.
 */
#include <complex.h>
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)

/* ---- double ---- */
#define OT double
#define OFN(name) OCAT(name, d)
#define OSQRT sqrt
#define OCOPYSIGN copysign
#include "oracle_generic.h"
#undef OT
#undef OFN
#undef OSQRT
#undef OCOPYSIGN

/* ---- float ---- */
#define OT float
#define OFN(name) OCAT(name, s)
#define OSQRT sqrtf
#define OCOPYSIGN copysignf
#include "oracle_generic.h"
#undef OT
#undef OFN
#undef OSQRT
#undef OCOPYSIGN

/* ---- __float128 (BigFloat stand-in) ---- */
#define OT __float128
#define OFN(name) OCAT(name, q)
#define OSQRT sqrtq
#define OCOPYSIGN copysignq
#include "oracle_generic.h"
#undef OT
#undef OFN
#undef OSQRT
#undef OCOPYSIGN

/* ---- complex double (suffix _z) and complex float (_c) ---- */
#define OC double _Complex
#define OR double
#define OFN(name) OCAT(name, z)
#define OCSQRT csqrt
#define OSQRTR sqrt
#define OCONJ conj
#define OCREAL creal
#define OCIMAG cimag
#define OCOPYSIGNR copysign
#include "oracle_complex.h"
#undef OC
#undef OR
#undef OFN
#undef OCSQRT
#undef OSQRTR
#undef OCONJ
#undef OCREAL
#undef OCIMAG
#undef OCOPYSIGNR

#define OC float _Complex
#define OR float
#define OFN(name) OCAT(name, c)
#define OCSQRT csqrtf
#define OSQRTR sqrtf
#define OCONJ conjf
#define OCREAL crealf
#define OCIMAG cimagf
#define OCOPYSIGNR copysignf
#include "oracle_complex.h"
#undef OC
#undef OR
#undef OFN
#undef OCSQRT
#undef OSQRTR
#undef OCONJ
#undef OCREAL
#undef OCIMAG
#undef OCOPYSIGNR

/* ---- double-in/double-out wrappers that compute in __float128 ---- */
static __float128 *q_from_d(int64_t m, int64_t n, const double *A, int64_t lda)
{
    __float128 *Q = (__float128 *)malloc(sizeof(__float128) * (size_t)(m * n > 0 ? m * n : 1));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) Q[i + j * m] = (__float128)A[i + j * lda];
    return Q;
}
static void d_from_q(int64_t m, int64_t n, const __float128 *Q, double *A, int64_t lda)
{
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) A[i + j * lda] = (double)Q[i + j * m];
}

int64_t orq_chol_recursive(int64_t n, double *A, int64_t lda, int64_t cutoff)
{
    __float128 *Q = q_from_d(n, n, A, lda);
    int64_t info = orc_chol_recursive_q(n, Q, n, cutoff);
    d_from_q(n, n, Q, A, lda);
    free(Q);
    return info;
}

int64_t orq_chol_unblocked(int64_t n, double *A, int64_t lda)
{
    __float128 *Q = q_from_d(n, n, A, lda);
    int64_t info = orc_chol_unblocked_q(n, Q, n);
    d_from_q(n, n, Q, A, lda);
    free(Q);
    return info;
}

void orq_qr_unblocked(int64_t m, int64_t n, double *A, int64_t lda, double *tau)
{
    __float128 *Q = q_from_d(m, n, A, lda);
    int64_t k = m < n ? m : n;
    __float128 *tq = (__float128 *)malloc(sizeof(__float128) * (size_t)(k > 0 ? k : 1));
    orc_qr_unblocked_q(m, n, Q, m, tq);
    d_from_q(m, n, Q, A, lda);
    for (int64_t i = 0; i < k; ++i) tau[i] = (double)tq[i];
    free(tq);
    free(Q);
}

/* Backward error of a Cholesky factor accumulated in __float128:
 * returns ‖A − L Lᵀ‖_F / ‖A‖_F using the lower triangle of A (symmetric)
 * and the lower triangle of L. */
double orq_chol_residual(int64_t n, const double *A, int64_t lda, const double *L, int64_t ldl)
{
    __float128 num = 0, den = 0;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            int64_t r = i > j ? i : j, c = i > j ? j : i;
            __float128 a = (__float128)A[r + c * lda];
            __float128 s = 0;
            int64_t kmax = i < j ? i : j;
            for (int64_t k = 0; k <= kmax; ++k)
                s += (__float128)L[i + k * ldl] * (__float128)L[j + k * ldl];
            __float128 d = a - s;
            num += d * d;
            den += a * a;
        }
    return (double)sqrtq(num / den);
}

/* symtri! / bidiagonalize! evaluated in __float128 from float64 input,
 * rounded back: the stand-in for the reference's BigFloat runs, the truth
 * the FP64 results are measured against. */
void orq_symtri(int upper, int64_t n, double *A, int64_t lda, double *tau)
{
    __float128 *Q = q_from_d(n, n, A, lda);
    __float128 *tq = (__float128 *)malloc(sizeof(__float128) * (size_t)(n > 1 ? n - 1 : 1));
    if (upper)
        orc_symtri_upper_q(n, Q, n, tq);
    else
        orc_symtri_lower_q(n, Q, n, tq);
    d_from_q(n, n, Q, A, lda);
    for (int64_t i = 0; i + 1 < n; ++i) tau[i] = (double)tq[i];
    free(tq);
    free(Q);
}

void orq_bidiag(int64_t m, int64_t n, double *A, int64_t lda, double *taul, double *taur)
{
    __float128 *Q = q_from_d(m, n, A, lda);
    int64_t k = m < n ? m : n;
    __float128 *tl = (__float128 *)malloc(sizeof(__float128) * (size_t)(k > 0 ? k : 1));
    __float128 *tr = (__float128 *)malloc(sizeof(__float128) * (size_t)(k > 0 ? k : 1));
    orc_bidiag_q(m, n, Q, m, tl, tr);
    d_from_q(m, n, Q, A, lda);
    int64_t nl = m >= n ? k : k - 1, nr = m >= n ? k - 1 : k;
    for (int64_t i = 0; i < nl; ++i) taul[i] = (double)tl[i];
    for (int64_t i = 0; i < nr; ++i) taur[i] = (double)tr[i];
    free(tl);
    free(tr);
    free(Q);
}
