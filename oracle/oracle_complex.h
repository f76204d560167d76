/*
 * Complex body of the CPU oracle (TEST INFRASTRUCTURE ONLY).  Included by
 * gla_oracle.c once per complex type with OC (the complex type), OR (its
 * real part type), OFN(name), OCSQRT, OSQRTR, OCONJ, OCREAL, OCIMAG,
 * OCOPYSIGNR defined.  Column-major, 0-based; each loop cites the 1-based
 * reference lines it restates.
 */

#define IXC(i, j, ld) ((int64_t)(i) + (int64_t)(j) * (int64_t)(ld))

/* rankUpdate!(A::Hermitian, a, α) for complex A — the BLAS her! the
 * reference dispatches to (src/juliaBLAS.jl:41-48), restated as its
 * generic loop (:53-63): A[i,j] += a[i] α a[j]' for i ≥ j, the diagonal's
 * imaginary part set to zero as her! leaves it. */
void OFN(orz_her_l)(int64_t n, OR alpha, const OC *a, OC *A, int64_t lda)
{
    for (int64_t j = 0; j < n; ++j) {
        OC ajc = OCONJ(a[j]);
        for (int64_t i = j; i < n; ++i) A[IXC(i, j, lda)] += a[i] * alpha * ajc;
        A[IXC(j, j, lda)] = OCREAL(A[IXC(j, j, lda)]);
    }
}

/* rankUpdate!(C::Hermitian, A, α) for complex C — BLAS herk!
 * (src/juliaBLAS.jl:78-86), as the generic k-outer loop (:89-112), the
 * diagonal kept real. */
void OFN(orz_herk_l)(int64_t n, int64_t k, OR alpha, const OC *A, int64_t lda, OC *C, int64_t ldc)
{
    for (int64_t kk = 0; kk < k; ++kk)
        for (int64_t j = 0; j < n; ++j) {
            OC ajkc = OCONJ(A[IXC(j, kk, lda)]);
            for (int64_t i = j; i < n; ++i) C[IXC(i, j, ldc)] += A[IXC(i, kk, lda)] * alpha * ajkc;
        }
    for (int64_t j = 0; j < n; ++j) C[IXC(j, j, ldc)] = OCREAL(C[IXC(j, j, ldc)]);
}

/* rdiv!(B, LowerTriangular(L)') for complex: X Lᴴ = B by row-wise forward
 * substitution, X[i,j] = (B[i,j] − Σ_{k<j} X[i,k] conj(L[j,k])) / conj(L[j,j]). */
void OFN(orz_trsm_rlc)(int64_t m, int64_t n, const OC *L, int64_t ldl, OC *B, int64_t ldb)
{
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            OC s = B[IXC(i, j, ldb)];
            for (int64_t k = 0; k < j; ++k) s -= B[IXC(i, k, ldb)] * OCONJ(L[IXC(j, k, ldl)]);
            B[IXC(i, j, ldb)] = s / OCONJ(L[IXC(j, j, ldl)]);
        }
}

/* cholUnblocked!(A, Val{:L}) for complex A, src/cholesky.jl:3-15: the
 * complex sqrt of the pivot (:5; a negative real pivot gives an imaginary
 * root and no error — returns the 1-based index of the first such pivot
 * for the caller's information), a21 scaled by inv(real(A[1,1])) (:8),
 * her! on A22 (:11). */
int64_t OFN(orz_chol_unblocked)(int64_t n, OC *A, int64_t lda)
{
    int64_t first_bad = 0;
    for (int64_t k = 0; k < n; ++k) {
        OC *Akk = A + IXC(k, k, lda);
        if (OCREAL(Akk[0]) < (OR)0 && !first_bad) first_bad = k + 1;
        Akk[0] = OCSQRT(Akk[0]);                                    /* :5 */
        if (k + 1 < n) {
            OR r = (OR)1 / OCREAL(Akk[0]);                          /* :8 */
            for (int64_t i = 1; i < n - k; ++i) Akk[i] *= r;
            OFN(orz_her_l)(n - k - 1, (OR)-1, Akk + 1, A + IXC(k + 1, k + 1, lda), lda); /* :11 */
        }
    }
    return first_bad;
}

/* cholBlocked!(A, Val{:L}, blocksize), src/cholesky.jl:17-35 (complex). */
int64_t OFN(orz_chol_blocked)(int64_t n, OC *A, int64_t lda, int64_t bs)
{
    int64_t off = 0, bad = 0;
    while (1) {
        int64_t nn = n - off;
        OC *A11 = A + IXC(off, off, lda);
        int64_t mnb = nn < bs ? nn : bs;                            /* :23 */
        int64_t b = OFN(orz_chol_unblocked)(mnb, A11, lda);         /* :25 */
        if (b && !bad) bad = off + b;
        if (!(nn > bs)) break;                                      /* :26 */
        OFN(orz_trsm_rlc)(nn - bs, bs, A11, lda, A11 + bs, lda);    /* :27-28 */
        OFN(orz_herk_l)(nn - bs, bs, (OR)-1, A11 + bs, lda, A11 + IXC(bs, bs, lda), lda); /* :30-31 */
        off += bs;                                                  /* :32 */
    }
    return bad;
}

/* cholRecursive!(A, Val{:L}, cutoff), src/cholesky.jl:37-55 (complex). */
int64_t OFN(orz_chol_recursive)(int64_t n, OC *A, int64_t lda, int64_t cutoff)
{
    if (n == 1) {                                                   /* :39-40 */
        int64_t bad = OCREAL(A[0]) < (OR)0;
        A[0] = OCSQRT(A[0]);
        return bad;
    }
    if (n < cutoff) return OFN(orz_chol_unblocked)(n, A, lda);      /* :41-42 */
    int64_t n2 = n / 2;                                             /* :44 */
    int64_t bad = OFN(orz_chol_recursive)(n2, A, lda, 1);           /* :46 */
    OFN(orz_trsm_rlc)(n - n2, n2, A, lda, A + n2, lda);             /* :47-48 */
    OFN(orz_herk_l)(n - n2, n2, (OR)-1, A + n2, lda, A + IXC(n2, n2, lda), lda); /* :50-51 */
    int64_t b2 = OFN(orz_chol_recursive)(n - n2, A + IXC(n2, n2, lda), lda, cutoff); /* :52 */
    if (!bad && b2) bad = n2 + b2;
    return bad;
}

/* LinearAlgebra.reflector!(x) for complex x (the stdlib, called at
 * src/qr.jl:96): ν = copysign(‖x‖, real(x₁)); ξ₁ = x₁ + ν; x₁ <- −ν;
 * x[2:n] ./= ξ₁; τ = ξ₁/ν; a zero norm gives τ = 0. */
OC OFN(orz_reflector)(int64_t n, OC *x)
{
    if (n == 0) return (OC)0;
    OC xi1 = x[0];
    OR s = 0;
    for (int64_t i = 0; i < n; ++i) s += OCREAL(x[i]) * OCREAL(x[i]) + OCIMAG(x[i]) * OCIMAG(x[i]);
    OR normu = OSQRTR(s);
    if (normu == (OR)0) return (OC)0;
    OR nu = OCOPYSIGNR(normu, OCREAL(xi1));
    xi1 += nu;
    x[0] = -nu;
    for (int64_t i = 1; i < n; ++i) x[i] /= xi1;
    return xi1 / nu;
}

/* LinearAlgebra.reflectorApply!(x, τ, A) for complex (the stdlib, left,
 * src/qr.jl:102): vAj = conj(τ) (A[1,j] + Σ_{i≥2} conj(x_i) A[i,j]);
 * A[1,j] −= vAj; A[i,j] −= x_i vAj. */
void OFN(orz_reflector_apply)(int64_t m, int64_t n, const OC *x, OC tau, OC *A, int64_t lda)
{
    if (m == 0) return;
    for (int64_t j = 0; j < n; ++j) {
        OC *Aj = A + IXC(0, j, lda);
        OC d = 0;
        for (int64_t i = 1; i < m; ++i) d += OCONJ(x[i]) * Aj[i];
        OC vAj = OCONJ(tau) * (Aj[0] + d);
        Aj[0] -= vAj;
        for (int64_t i = 1; i < m; ++i) Aj[i] -= x[i] * vAj;
    }
}

/* qrUnblocked!(A, τ) for complex, src/qr.jl:86-111. */
void OFN(orz_qr_unblocked)(int64_t m, int64_t n, OC *A, int64_t lda, OC *tau)
{
    int64_t j = 0;
    while (1) {
        int64_t mm = m - j, nn = n - j;
        OC *a = A + IXC(j, j, lda);
        OC t1 = OFN(orz_reflector)(mm, a);                                 /* :96 */
        tau[j] = t1;                                                       /* :98 */
        if (nn > 1) OFN(orz_reflector_apply)(mm, nn - 1, a, t1, a + lda, lda); /* :101-103 */
        if (!(mm > 1 && nn > 1)) break;                                    /* :106-108 */
        ++j;
    }
}

#undef IXC
