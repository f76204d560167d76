/*
 * Generic body of the CPU oracle.  Included once per element type by
 * gla_oracle.c with OT (the element type), OFN(name) (the suffixed symbol
 * name), OSQRT, OCOPYSIGN defined.  TEST INFRASTRUCTURE ONLY: nothing in the
 * product path links this file.
 *
 * Storage is column-major with an explicit leading dimension, as in Julia.
 * Indices are 0-based here; every loop cites the 1-based reference lines it
 * restates.
 */

#define IX(i, j, ld) ((int64_t)(i) + (int64_t)(j) * (int64_t)(ld))

/* rankUpdate!(A::Hermitian, a::StridedVector, α::Real), uplo = 'L'.
 * src/juliaBLAS.jl:53-63 — j outer, i = j:n inner, A[i,j] += a[i]*α*a[j]'. */
void OFN(orc_syr_l)(int64_t n, OT alpha, const OT *a, int64_t inca, OT *A, int64_t lda)
{
    for (int64_t j = 0; j < n; ++j) {
        OT ajc = a[j * inca];
        for (int64_t i = j; i < n; ++i)
            A[IX(i, j, lda)] += a[i * inca] * alpha * ajc;
    }
}

/* rankUpdate!(C::Hermitian, A::StridedVecOrMat, α::Real), uplo = 'L'.
 * src/juliaBLAS.jl:89-112 — k outer, j middle, i = j:n inner. */
void OFN(orc_syrk_l)(int64_t n, int64_t k, OT alpha, const OT *A, int64_t lda, OT *C,
                     int64_t ldc)
{
    for (int64_t kk = 0; kk < k; ++kk)
        for (int64_t j = 0; j < n; ++j) {
            OT ajkc = A[IX(j, kk, lda)];
            for (int64_t i = j; i < n; ++i)
                C[IX(i, j, ldc)] += A[IX(i, kk, lda)] * alpha * ajkc;
        }
}

/* rdiv!(B, LowerTriangular(L)') : B <- B * L^{-T}, i.e. solve X*L^T = B.
 * Called at src/cholesky.jl:28,48.  Restated as the stdlib's generic
 * row-wise forward substitution with U = L^T: X[i,j] = (B[i,j] - Σ_{k<j}
 * X[i,k] U[k,j]) / U[j,j], U[k,j] = L[j,k]. */
void OFN(orc_trsm_rlt)(int64_t m, int64_t n, const OT *L, int64_t ldl, OT *B, int64_t ldb)
{
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            OT s = B[IX(i, j, ldb)];
            for (int64_t k = 0; k < j; ++k)
                s -= B[IX(i, k, ldb)] * L[IX(j, k, ldl)];
            B[IX(i, j, ldb)] = s / L[IX(j, j, ldl)];
        }
}

/* sqrt with Julia's DomainError: returns 0 and sets *bad if x < 0. */
static inline OT OFN(orc_sqrt_checked)(OT x, int *bad)
{
    if (x < (OT)0) {
        *bad = 1;
        return (OT)0;
    }
    return OSQRT(x);
}

/* cholUnblocked!(A, Val{:L}) src/cholesky.jl:3-15.
 * Returns 0, or the 1-based index (global within this call) of the pivot
 * whose square root would throw DomainError; A is then partially updated
 * as in the reference. */
int64_t OFN(orc_chol_unblocked)(int64_t n, OT *A, int64_t lda)
{
    /* the tail recursion on A22 (:12) is unrolled into a loop over k */
    for (int64_t k = 0; k < n; ++k) {
        int bad = 0;
        OT *Akk = A + IX(k, k, lda);
        Akk[0] = OFN(orc_sqrt_checked)(Akk[0], &bad);          /* :5 */
        if (bad) return k + 1;
        if (k + 1 < n) {
            OT r = (OT)1 / Akk[0];                             /* :8 inv(real(A11)) */
            for (int64_t i = 1; i < n - k; ++i) Akk[i] *= r;   /* :8 rmul! */
            /* :11 rankUpdate!(Hermitian(A22,:L), a21, -1) */
            OFN(orc_syr_l)(n - k - 1, (OT)-1, Akk + 1, 1, A + IX(k + 1, k + 1, lda), lda);
        }
    }
    return 0;
}

/* cholBlocked!(A, Val{:L}, blocksize) src/cholesky.jl:17-35. */
int64_t OFN(orc_chol_blocked)(int64_t n, OT *A, int64_t lda, int64_t bs)
{
    int64_t off = 0;                      /* tail recursion (:32) as a loop */
    while (1) {
        int64_t nn = n - off;
        OT *Ab = A + IX(off, off, lda);
        int64_t mnb = nn < bs ? nn : bs;                       /* :23 */
        int64_t info = OFN(orc_chol_unblocked)(mnb, Ab, lda);  /* :25 */
        if (info) return off + info;
        if (nn <= bs) break;                                   /* :26 */
        /* :27-28 rdiv!(A21, LowerTriangular(A11)') */
        OFN(orc_trsm_rlt)(nn - bs, bs, Ab, lda, Ab + bs, lda);
        /* :30-31 rankUpdate!(Hermitian(A22,:L), A21, -1) */
        OFN(orc_syrk_l)(nn - bs, bs, (OT)-1, Ab + bs, lda, Ab + IX(bs, bs, lda), lda);
        off += bs;
    }
    return 0;
}

/* cholRecursive!(A, Val{:L}, cutoff) src/cholesky.jl:37-55.  Note the
 * reference does not forward cutoff to the A11 recursion (:46). */
int64_t OFN(orc_chol_recursive)(int64_t n, OT *A, int64_t lda, int64_t cutoff)
{
    if (n == 1) {                                              /* :39-40 */
        int bad = 0;
        A[0] = OFN(orc_sqrt_checked)(A[0], &bad);
        return bad ? 1 : 0;
    }
    if (n < cutoff) return OFN(orc_chol_unblocked)(n, A, lda); /* :41-42 */
    int64_t n2 = n / 2;                                        /* :44 */
    int64_t info = OFN(orc_chol_recursive)(n2, A, lda, 1);     /* :46 */
    if (info) return info;
    OFN(orc_trsm_rlt)(n - n2, n2, A, lda, A + n2, lda);        /* :47-48 */
    OFN(orc_syrk_l)(n - n2, n2, (OT)-1, A + n2, lda, A + IX(n2, n2, lda), lda); /* :50-51 */
    info = OFN(orc_chol_recursive)(n - n2, A + IX(n2, n2, lda), lda, cutoff);  /* :52 */
    return info ? n2 + info : 0;
}

/* LinearAlgebra.reflector!(x) (Julia stdlib, called at src/qr.jl:96):
 * ν = copysign(‖x‖, x1); ξ1 = x1 + ν; x1 <- -ν; x[2:n] ./= ξ1; τ = ξ1/ν;
 * zero norm gives τ = 0 and leaves x alone.  n == 0 gives τ = 0. */
OT OFN(orc_reflector)(int64_t n, OT *x)
{
    if (n == 0) return (OT)0;
    OT xi1 = x[0];
    OT s = 0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    OT normu = OSQRT(s);
    if (normu == (OT)0) return (OT)0;
    OT nu = OCOPYSIGN(normu, xi1);
    xi1 += nu;
    x[0] = -nu;
    for (int64_t i = 1; i < n; ++i) x[i] /= xi1;
    return xi1 / nu;
}

/* LinearAlgebra.reflectorApply!(x, τ, A) (Julia stdlib, left application,
 * called at src/qr.jl:102): for each column j,
 * vAj = τ*(A[1,j] + dot(x[2:m], A[2:m,j])); A[1,j] -= vAj; A[2:m,j] -= vAj*x[2:m]. */
void OFN(orc_reflector_apply)(int64_t m, int64_t n, const OT *x, OT tau, OT *A, int64_t lda)
{
    if (m == 0) return;
    for (int64_t j = 0; j < n; ++j) {
        OT *Aj = A + IX(0, j, lda);
        OT d = 0;
        for (int64_t i = 1; i < m; ++i) d += x[i] * Aj[i];
        OT vAj = tau * (Aj[0] + d);
        Aj[0] -= vAj;
        for (int64_t i = 1; i < m; ++i) Aj[i] -= vAj * x[i];
    }
}

/* qrUnblocked!(A, τ) src/qr.jl:86-111 (tail recursion unrolled). */
void OFN(orc_qr_unblocked)(int64_t m, int64_t n, OT *A, int64_t lda, OT *tau)
{
    int64_t j = 0;
    while (1) {
        int64_t mm = m - j, nn = n - j;
        OT *a = A + IX(j, j, lda);
        OT t1 = OFN(orc_reflector)(mm, a);                                 /* :96 */
        tau[j] = t1;                                                       /* :98 */
        if (nn > 1) OFN(orc_reflector_apply)(mm, nn - 1, a, t1, a + lda, lda); /* :101-103 */
        if (!(mm > 1 && nn > 1)) break;                                    /* :106-108 */
        ++j;
    }
}

/* getindex(::QR2, Tuple{:QBlocked}) src/qr.jl:64-83 — the UT-transform
 * T = (I + D striu(VᴴV))⁻¹ D written into Tm (k×k, k = min(m,n)). */
void OFN(orc_qr_tblock)(int64_t m, int64_t n, const OT *F, int64_t ldf, const OT *tau, OT *Tm,
                        int64_t ldt)
{
    int64_t k = m < n ? m : n;
    for (int64_t j = 0; j < k; ++j)
        for (int64_t i = 0; i < k; ++i) Tm[IX(i, j, ldt)] = 0;
    for (int64_t j = 0; j < k; ++j)                                        /* :69-74 */
        for (int64_t i = 0; i < j; ++i) {
            OT d = 0;
            for (int64_t r = j + 1; r < m; ++r) d += F[IX(r, i, ldf)] * F[IX(r, j, ldf)];
            Tm[IX(i, j, ldt)] = tau[i] * (F[IX(j, i, ldf)] + d);
        }
    /* :75 inv!(UnitUpperTriangular(Tmat)), restated as the column sweep of
     * LAPACK dtrti2('U','U'): column j <- -(inverse so far) * column j. */
    for (int64_t j = 0; j < k; ++j) {
        for (int64_t i = 0; i < j; ++i) {          /* x = triu_unit(Tinv[0:j,0:j]) * x */
            OT s = Tm[IX(i, j, ldt)];
            for (int64_t l = i + 1; l < j; ++l) s += Tm[IX(i, l, ldt)] * Tm[IX(l, j, ldt)];
            Tm[IX(i, j, ldt)] = s;
        }
        for (int64_t i = 0; i < j; ++i) Tm[IX(i, j, ldt)] = -Tm[IX(i, j, ldt)];
    }
    for (int64_t j = 0; j < k; ++j) {                                      /* :76-81 */
        Tm[IX(j, j, ldt)] = tau[j];
        for (int64_t i = 0; i < j; ++i) Tm[IX(i, j, ldt)] *= tau[j];
    }
}

/* lmul!(A::UnitLowerTriangular, B, α) src/juliaBLAS.jl:154-166:
 * B <- α * L_unit * B, rows processed bottom-up. */
void OFN(orc_lmul_unit_lower)(int64_t m, int64_t n, const OT *L, int64_t ldl, OT *B,
                              int64_t ldb, OT alpha)
{
    for (int64_t i = m - 1; i >= 0; --i)
        for (int64_t j = 0; j < n; ++j) {
            OT s = alpha * B[IX(i, j, ldb)];
            for (int64_t l = 0; l < i; ++l) s += alpha * L[IX(i, l, ldl)] * B[IX(l, j, ldb)];
            B[IX(i, j, ldb)] = s;
        }
}

/* lmul!(H::HouseholderBlock, A, M) (trans = 0) and
 * lmul!(adjoint(H), A, M) (trans = 1), src/householder.jl:82-115,119-157.
 * V is mA×nV (nV = size(V,2)), T is bsz×bsz upper, bsz = min(mA, nV);
 * M is bsz×nA workspace. */
void OFN(orc_wy_apply)(int trans, int64_t mA, int64_t nA, const OT *V, int64_t ldv,
                       int64_t nV, const OT *Tm, int64_t ldt, OT *A, int64_t lda, OT *M,
                       int64_t ldm)
{
    int64_t b = mA < nV ? mA : nV;                                  /* :124-125 */
    /* :140 copyto!(M, A1) */
    for (int64_t j = 0; j < nA; ++j)
        for (int64_t i = 0; i < b; ++i) M[IX(i, j, ldm)] = A[IX(i, j, lda)];
    /* :142 lmul!(V1', M): M <- V1ᵀ M, V1 unit lower; row i uses old rows l > i. */
    for (int64_t j = 0; j < nA; ++j)
        for (int64_t i = 0; i < b; ++i) {
            OT s = M[IX(i, j, ldm)];
            for (int64_t l = i + 1; l < b; ++l) s += V[IX(l, i, ldv)] * M[IX(l, j, ldm)];
            M[IX(i, j, ldm)] = s;
        }
    /* :144 mul!(M, V2', A2, 1, 1) */
    for (int64_t j = 0; j < nA; ++j)
        for (int64_t i = 0; i < b; ++i) {
            OT s = 0;
            for (int64_t r = b; r < mA; ++r) s += V[IX(r, i, ldv)] * A[IX(r, j, lda)];
            M[IX(i, j, ldm)] += s;
        }
    /* :146 lmul!(T', M) (adjoint) or :104 lmul!(T, M) */
    for (int64_t j = 0; j < nA; ++j) {
        if (trans) {
            for (int64_t i = b - 1; i >= 0; --i) {      /* Tᵀ lower: uses old rows l <= i */
                OT s = 0;
                for (int64_t l = 0; l <= i; ++l) s += Tm[IX(l, i, ldt)] * M[IX(l, j, ldm)];
                M[IX(i, j, ldm)] = s;
            }
        } else {
            for (int64_t i = 0; i < b; ++i) {           /* T upper: uses old rows l >= i */
                OT s = 0;
                for (int64_t l = i; l < b; ++l) s += Tm[IX(i, l, ldt)] * M[IX(l, j, ldm)];
                M[IX(i, j, ldm)] = s;
            }
        }
    }
    /* :149 mul!(A2, V2, M, -1, 1) */
    for (int64_t j = 0; j < nA; ++j)
        for (int64_t r = b; r < mA; ++r) {
            OT s = 0;
            for (int64_t i = 0; i < b; ++i) s += V[IX(r, i, ldv)] * M[IX(i, j, ldm)];
            A[IX(r, j, lda)] -= s;
        }
    /* :153 lmul!(V1, M, -1) (juliaBLAS.jl:154) ; :154 axpy!(1, M, A1) */
    OFN(orc_lmul_unit_lower)(b, nA, V, ldv, M, ldm, (OT)-1);
    for (int64_t j = 0; j < nA; ++j)
        for (int64_t i = 0; i < b; ++i) A[IX(i, j, lda)] += M[IX(i, j, ldm)];
}

/* qrBlocked!(A, blocksize, τ, work) src/qr.jl:113-146 (tail recursion as a
 * loop; work is blocksize × n, ldw >= blocksize). Scratch for T is malloc'd. */
void OFN(orc_qr_blocked)(int64_t m, int64_t n, OT *A, int64_t lda, int64_t bs, OT *tau,
                         OT *work, int64_t ldw)
{
    int64_t off = 0;
    OT *Tm = (OT *)malloc(sizeof(OT) * (size_t)(bs * bs > 0 ? bs * bs : 1));
    while (1) {
        int64_t mm = m - off, nn = n - off;
        OT *Ab = A + IX(off, off, lda);
        int64_t w = nn < bs ? nn : bs;                                     /* :121 */
        OFN(orc_qr_unblocked)(mm, w, Ab, lda, tau + off);                  /* :123 */
        if (nn > bs) {                                                     /* :126 */
            /* F[Tuple{:QBlocked}] of the panel QR2 (mm × w), :132 */
            int64_t k = mm < w ? mm : w;
            OFN(orc_qr_tblock)(mm, w, Ab, lda, tau + off, Tm, k);
            /* :132 lmul!(F', A2, view(work, 1:min(m,bs), 1:n-bs)) */
            OFN(orc_wy_apply)(1, mm, nn - bs, Ab, lda, w, Tm, k, Ab + IX(0, bs, lda), lda,
                              work, ldw);
        }
        if (!(mm > bs && nn > bs)) break;                                  /* :136 */
        off += bs;
    }
    free(Tm);
}

/* ------------------------------------------------------------------ LDLᵀ -- */
/* _ldlt_lower!(A) src/ldlt.jl:16-33: column 1 is divided by δ = A[1,1]
 * (:19-22), the lower trailing block takes −l_i (l_j δ)' (:23-28), then the
 * recursion on view(A, 2:n, 2:n) while n > 2 (:29-31), unrolled into k. */
void OFN(orc_ldlt_lower)(int64_t n, OT *A, int64_t lda)
{
    for (int64_t k = 0; k + 1 < n; ++k) {
        OT d = A[IX(k, k, lda)];
        for (int64_t i = k + 1; i < n; ++i) A[IX(i, k, lda)] = A[IX(i, k, lda)] / d;   /* :19-22 */
        for (int64_t j = k + 1; j < n; ++j) {                                             /* :23-28 */
            OT ljd = A[IX(j, k, lda)] * d;
            for (int64_t i = j; i < n; ++i) A[IX(i, j, lda)] -= A[IX(i, k, lda)] * ljd;
        }
    }
}

/* _ldlt_upper!(A) src/ldlt.jl:48-62: row 1 right of the diagonal takes
 * A[1,j] / δ, each column j the update A[i,j] −= A[1,i]' (δ u_j) for
 * i = 2:j with the undivided A[1,j] (:51-56); recursion while n > 2. */
void OFN(orc_ldlt_upper)(int64_t n, OT *A, int64_t lda)
{
    for (int64_t k = 0; k + 1 < n; ++k) {
        OT d = A[IX(k, k, lda)];
        for (int64_t j = k + 1; j < n; ++j) {
            OT dl = A[IX(k, j, lda)];
            A[IX(k, j, lda)] = dl / d;
            for (int64_t i = k + 1; i <= j; ++i) A[IX(i, j, lda)] -= A[IX(k, i, lda)] * dl;
        }
    }
}

/* _ldlt_lower_blocked!(A, blocksize) src/ldlt.jl:80-104 (tail recursion
 * on the trailing block unrolled).  rdiv! by UnitLowerTriangular(A11)' is
 * the row-wise substitution X[i,j] = B[i,j] − Σ_{l<j} X[i,l] A11[j,l]
 * (:86), rdiv! by Diagonal(d₁) a division per column (:87). */
void OFN(orc_ldlt_lower_blocked)(int64_t n, OT *A, int64_t lda, int64_t bs)
{
    OT *work = (OT *)malloc(sizeof(OT) * (size_t)(bs > 0 ? bs : 1));
    int64_t k = 0;
    while (bs < n - k) {                                        /* :82 */
        OT *A11 = A + IX(k, k, lda), *A21 = A11 + bs, *A22 = A11 + IX(bs, bs, lda);
        int64_t m = n - k - bs;
        OFN(orc_ldlt_lower)(bs, A11, lda);                      /* :84 */
        for (int64_t i = 0; i < m; ++i)                         /* :86 */
            for (int64_t j = 0; j < bs; ++j) {
                OT s = A21[IX(i, j, lda)];
                for (int64_t l = 0; l < j; ++l) s -= A21[IX(i, l, lda)] * A11[IX(j, l, lda)];
                A21[IX(i, j, lda)] = s;
            }
        for (int64_t j = 0; j < bs; ++j)                        /* :87 */
            for (int64_t i = 0; i < m; ++i) A21[IX(i, j, lda)] = A21[IX(i, j, lda)] / A11[IX(j, j, lda)];
        for (int64_t j = 0; j < m; ++j) {                       /* :88-95 */
            for (int64_t l = 0; l < bs; ++l) work[l] = A11[IX(l, l, lda)] * A21[IX(j, l, lda)];
            for (int64_t i = j; i < m; ++i)
                for (int64_t l = 0; l < bs; ++l) A22[IX(i, j, lda)] -= A21[IX(i, l, lda)] * work[l];
        }
        k += bs;                                                /* :96 */
    }
    OFN(orc_ldlt_lower)(n - k, A + IX(k, k, lda), lda);         /* :101 */
    free(work);
}

/* _ldlt_upper_blocked!(A, blocksize) src/ldlt.jl:121-148: U₁₂ <- D₁⁻¹
 * (U₁₁')⁻¹ A₁₂ by column-wise forward substitution with the unit lower
 * U₁₁' (:126) and a division per row (:127), then the upper trailing
 * update A22[i,j] −= Σ_l U₁₂[l,i]' (d_l U₁₂[l,j]) for i ≤ j (:128-135). */
void OFN(orc_ldlt_upper_blocked)(int64_t n, OT *A, int64_t lda, int64_t bs)
{
    OT *work = (OT *)malloc(sizeof(OT) * (size_t)(bs > 0 ? bs : 1));
    int64_t k = 0;
    while (bs < n - k) {                                        /* :123 */
        OT *A11 = A + IX(k, k, lda), *U12 = A11 + IX(0, bs, lda), *A22 = A11 + IX(bs, bs, lda);
        int64_t m = n - k - bs;
        OFN(orc_ldlt_upper)(bs, A11, lda);                      /* :125 */
        for (int64_t j = 0; j < m; ++j)                         /* :126 */
            for (int64_t i = 0; i < bs; ++i) {
                OT s = U12[IX(i, j, lda)];
                for (int64_t l = 0; l < i; ++l) s -= A11[IX(l, i, lda)] * U12[IX(l, j, lda)];
                U12[IX(i, j, lda)] = s;
            }
        for (int64_t j = 0; j < m; ++j)                         /* :127 */
            for (int64_t i = 0; i < bs; ++i) U12[IX(i, j, lda)] = U12[IX(i, j, lda)] / A11[IX(i, i, lda)];
        for (int64_t j = 0; j < m; ++j) {                       /* :128-135 */
            for (int64_t l = 0; l < bs; ++l) work[l] = A11[IX(l, l, lda)] * U12[IX(l, j, lda)];
            for (int64_t i = 0; i <= j; ++i)
                for (int64_t l = 0; l < bs; ++l) A22[IX(i, j, lda)] -= U12[IX(l, i, lda)] * work[l];
        }
        k += bs;
    }
    OFN(orc_ldlt_upper)(n - k, A + IX(k, k, lda), lda);         /* :142 */
    free(work);
}

/* ldiv!(F::LDLt, X) src/ldlt.jl:225-228: X <- L⁻ᴴ D⁻¹ L⁻¹ X with L unit
 * lower (uplo 'L') or L = U' (uplo 'U'); the substitutions are the stdlib's
 * column-oriented ones, D a division per row. */
void OFN(orc_ldlt_solve)(int upper, int64_t n, int64_t nrhs, const OT *F, int64_t ldf, OT *B,
                         int64_t ldb)
{
    for (int64_t c = 0; c < nrhs; ++c) {
        OT *x = B + IX(0, c, ldb);
        for (int64_t j = 0; j < n; ++j)                 /* ldiv!(L, X): L[i,j] = F[i,j] or F[j,i] */
            for (int64_t i = j + 1; i < n; ++i)
                x[i] -= (upper ? F[IX(j, i, ldf)] : F[IX(i, j, ldf)]) * x[j];
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] / F[IX(i, i, ldf)];   /* ldiv!(D, X) */
        for (int64_t j = n - 1; j >= 0; --j)            /* ldiv!(L', X) */
            for (int64_t i = 0; i < j; ++i)
                x[i] -= (upper ? F[IX(i, j, ldf)] : F[IX(j, i, ldf)]) * x[j];
    }
}

/* ---------------------------------------------------- two-sided reductions -- */
/* LinearAlgebra.reflector! on the strided vector x[0], x[inc], …,
 * x[(n−1)·inc] (inc < 0 walks a column upwards, as view(A, m:-1:1, j)
 * does); the arithmetic of orc_reflector. */
OT OFN(orc_reflector_inc)(int64_t n, OT *x, int64_t inc)
{
    if (n == 0) return (OT)0;
    OT xi1 = x[0];
    OT s = 0;
    for (int64_t i = 0; i < n; ++i) s += x[i * inc] * x[i * inc];
    OT normu = OSQRT(s);
    if (normu == (OT)0) return (OT)0;
    OT nu = OCOPYSIGN(normu, xi1);
    xi1 += nu;
    x[0] = -nu;
    for (int64_t i = 1; i < n; ++i) x[i * inc] /= xi1;
    return xi1 / nu;
}

/* reflectorApply!(A, x, τ) from the right, the package's own method
 * (src/qr.jl:18-40, real case), called at src/svd.jl:344,363: for each
 * row i, s = τ (A[i,1] + Σ_{j≥2} A[i,j] x_j); A[i,1] −= s;
 * A[i,j] −= s x_j (x_j' = x_j for real x).  x is strided (incx). */
void OFN(orc_reflector_apply_right)(int64_t m, int64_t n, const OT *x, int64_t incx, OT tau,
                                    OT *A, int64_t lda)
{
    if (n == 0) return;
    for (int64_t i = 0; i < m; ++i) {
        OT s = A[IX(i, 0, lda)];
        for (int64_t j = 1; j < n; ++j) s += A[IX(i, j, lda)] * x[j * incx];
        s = s * tau;
        A[IX(i, 0, lda)] -= s;
        for (int64_t j = 1; j < n; ++j) A[IX(i, j, lda)] -= s * x[j * incx];
    }
}

/* symtriLower!(AS, τ, u) src/eigenSelfAdjoint.jl:450-504, real AS (n−2
 * reflectors; τ has n−1 entries, the last left 0).  Per column k: the
 * reflector of AS[k+1:n, k] (:466), ũ = τ·Hermitian(Ã,:L)·x (:477, the
 * symmetric product from the lower triangle), ξ = τ xᵀũ (:480), and the
 * lower triangle of Ã += x ξ xᵀ − x ũᵀ − ũ xᵀ (:484-491). */
void OFN(orc_symtri_lower)(int64_t n, OT *A, int64_t lda, OT *tau)
{
    OT *u = (OT *)malloc(sizeof(OT) * (size_t)(n > 0 ? n : 1));
    for (int64_t k = 0; k + 1 < n; ++k) tau[k] = 0;
    for (int64_t k = 0; k + 2 < n; ++k) {                        /* :462 */
        const int64_t m = n - k - 1;
        OT *x = A + IX(k + 1, k, lda);
        OT *At = A + IX(k + 1, k + 1, lda);
        OT tk = OFN(orc_reflector)(m, x);                        /* :466 */
        tau[k] = tk;                                             /* :467 */
        OT tmp = x[0];                                           /* :470-471 */
        x[0] = 1;
        for (int64_t i = 0; i < m; ++i) {                        /* :477 */
            OT sacc = 0;
            for (int64_t j = 0; j < m; ++j)
                sacc += (i >= j ? At[IX(i, j, lda)] : At[IX(j, i, lda)]) * x[j];
            u[i] = tk * sacc;
        }
        OT xi = 0;                                               /* :480 */
        for (int64_t i = 0; i < m; ++i) xi += x[i] * u[i];
        xi = tk * xi;
        for (int64_t j = 0; j < m; ++j) {                        /* :484-491 */
            OT xj = x[j], uj = u[j], xixj = xi * xj;
            for (int64_t i = j; i < m; ++i)
                At[IX(i, j, lda)] += x[i] * xixj - x[i] * uj - u[i] * xj;
        }
        x[0] = tmp;                                              /* :494 */
    }
    free(u);
}

/* symtriUpper!(AS, τ, u) src/eigenSelfAdjoint.jl:507-564, real AS: step k
 * reflects column n−k+1 above the diagonal walked upwards (:523-527, the
 * reflector's first element is the superdiagonal entry), with the
 * symmetric product from the upper triangle (:537) and the update of the
 * upper triangle of Ã = AS[1:n−k, 1:n−k] (:544-551). */
void OFN(orc_symtri_upper)(int64_t n, OT *A, int64_t lda, OT *tau)
{
    OT *u = (OT *)malloc(sizeof(OT) * (size_t)(n > 0 ? n : 1));
    for (int64_t k = 0; k + 1 < n; ++k) tau[k] = 0;
    for (int64_t k = 0; k + 2 < n; ++k) {
        const int64_t m = n - k - 1;                             /* n − k (1-based k) */
        OT *x = A + IX(0, m, lda);                               /* :523 */
        OT tk = OFN(orc_reflector_inc)(m, x + (m - 1), -1);      /* :524-527 */
        tau[k] = tk;
        OT tmp = x[m - 1];                                       /* :530-531 */
        x[m - 1] = 1;
        for (int64_t i = 0; i < m; ++i) {                        /* :537 */
            OT sacc = 0;
            for (int64_t j = 0; j < m; ++j)
                sacc += (i <= j ? A[IX(i, j, lda)] : A[IX(j, i, lda)]) * x[j];
            u[i] = tk * sacc;
        }
        OT xi = 0;                                               /* :540 */
        for (int64_t i = 0; i < m; ++i) xi += x[i] * u[i];
        xi = tk * xi;
        for (int64_t j = 0; j < m; ++j) {                        /* :544-551 */
            OT xj = x[j], uj = u[j], xixj = xi * xj;
            for (int64_t i = 0; i <= j; ++i)
                A[IX(i, j, lda)] += x[i] * xixj - x[i] * uj - u[i] * xj;
        }
        x[m - 1] = tmp;                                          /* :554 */
    }
    free(u);
}

/* bidiagonalize!(A) src/svd.jl:328-381, real A.  m ≥ n (upper bidiagonal,
 * :332-349): for each column i the left reflector of A[i:m, i] applied to
 * A[i:m, i+1:n], then (i < n) the right reflector of row i, A[i, i+1:n],
 * applied to A[i+1:m, i+1:n]; taul has n entries, taur n−1.  m < n (lower
 * bidiagonal, :356-375): the right reflector of A[i, i:n] first, then the
 * left one of A[i+1:m, i]; taur has m entries, taul m−1. */
void OFN(orc_bidiag)(int64_t m, int64_t n, OT *A, int64_t lda, OT *taul, OT *taur)
{
    if (m >= n) {
        for (int64_t i = 0; i < n; ++i) {
            OT *x = A + IX(i, i, lda);
            OT t = OFN(orc_reflector)(m - i, x);                                 /* :335-336 */
            taul[i] = t;                                                         /* :337 */
            OFN(orc_reflector_apply)(m - i, n - i - 1, x, t, A + IX(i, i + 1, lda), lda); /* :338 */
            if (i + 1 < n) {                                                     /* :339 */
                OT *y = A + IX(i, i + 1, lda);
                t = OFN(orc_reflector_inc)(n - i - 1, y, lda);                   /* :340-342 */
                taur[i] = t;                                                     /* :343 */
                OFN(orc_reflector_apply_right)(m - i - 1, n - i - 1, y, lda, t,
                                               A + IX(i + 1, i + 1, lda), lda);  /* :344 */
            }
        }
    } else {
        for (int64_t i = 0; i < m; ++i) {
            OT *y = A + IX(i, i, lda);
            OT t = OFN(orc_reflector_inc)(n - i, y, lda);                        /* :359-361 */
            taur[i] = t;                                                         /* :362 */
            OFN(orc_reflector_apply_right)(m - i - 1, n - i, y, lda, t, A + IX(i + 1, i, lda),
                                           lda);                                 /* :363 */
            if (i + 1 < m) {                                                     /* :364 */
                OT *x = A + IX(i + 1, i, lda);
                t = OFN(orc_reflector)(m - i - 1, x);                            /* :365-366 */
                taul[i] = t;                                                     /* :367 */
                OFN(orc_reflector_apply)(m - i - 1, n - i - 1, x, t, A + IX(i + 1, i + 1, lda),
                                         lda);                                   /* :368 */
            }
        }
    }
}

#undef IX
