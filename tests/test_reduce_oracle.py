"""Pins the oracle's two-sided Householder reductions on the CPU:

symtriLower!/symtriUpper! (src/eigenSelfAdjoint.jl:450-564):
- LAPACK dsytrd uses the same reflector convention (β = −sign(x₁)‖x‖,
  v₁ = 1, same packing), so d, e, τ and the stored reflectors agree to
  rounding (the upper case's τ in LAPACK's reverse order);
- the reference's own test (test/eigenselfadjoint.jl:61-67): Q orthogonal
  and Qᵀ A Q = T;
- T's eigenvalues are A's (LAPACK dsyevd).

bidiagonalize! (src/svd.jl:328-381), tall and wide:
- Qlᵀ A Qr = B with Ql, Qr orthogonal (the reflector products);
- B's singular values are A's (LAPACK dgesdd; the reference's test at
  test/svd.jl:100-112 checks the same);
- B's entries, up to sign, are those of an independent Golub–Kahan–Lanczos
  bidiagonalisation started from e₁ (full reorthogonalisation), which the
  implicit-Q theorem makes unique; the first left reflector is QR's first
  (LAPACK dgeqrf).
A planted slip (the rank-2 update without the ξ term) fails Qᵀ A Q = T.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.linalg.lapack as lp

import oracle as O
from conftest import eps_of, rel


def sym(n, rng, dtype=np.float64):
    G = rng.standard_normal((n, n))
    return (G + G.T).astype(dtype)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 64, 129])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symtri_matches_lapack_dsytrd(n, uplo, rng):
    A = sym(n, rng)
    AS, tau = O.symtri(A, uplo)
    d, e = O.symtri_diagonals(AS, uplo)
    c, dl, el, tl, info = lp.dsytrd(A, lower=1 if uplo == "L" else 0)
    assert info == 0
    tol = 50 * n * eps_of(np.float64)
    assert rel(d, dl) <= tol and rel(e, el) <= tol
    if n > 2:
        tref = tl[:n - 2] if uplo == "L" else tl[1:n - 1][::-1]
        assert rel(tau[:n - 2], tref) <= tol
        if uplo == "L":
            assert rel(np.tril(AS, -2), np.tril(c, -2)) <= tol
        else:
            assert rel(np.triu(AS, 2), np.triu(c, 2)) <= tol
    if n > 1:
        assert tau[-1] == 0


@pytest.mark.parametrize("n", [5, 40, 100])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symtri_reference_test(n, uplo, rng):
    """test/eigenselfadjoint.jl:61-67: Q'Q ≈ I and Q' A Q ≈ T."""
    A = sym(n, rng)
    AS, tau = O.symtri(A, uplo)
    Q = O.symtri_q(AS, tau, uplo)
    d, e = O.symtri_diagonals(AS, uplo)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 10 * n * eps_of(np.float64)
    assert rel(Q.T @ A @ Q, T) <= 10 * n * eps_of(np.float64)
    assert rel(np.linalg.eigvalsh(T), np.linalg.eigvalsh(A)) <= 10 * n * eps_of(np.float64)
    # only the uplo triangle is read: the other one may hold anything
    B = A.copy()
    B[np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)] = 7.0
    AS2, tau2 = O.symtri(B, uplo)
    tri = np.tril if uplo == "L" else np.triu
    assert np.array_equal(tri(AS2), tri(AS)) and np.array_equal(tau2, tau)


def test_symtri_float32(rng):
    A = sym(60, rng, np.float32)
    AS, tau = O.symtri(A, "L")
    c, dl, el, tl, info = lp.ssytrd(A, lower=1)
    d, e = O.symtri_diagonals(AS, "L")
    assert rel(d, dl) <= 50 * 60 * eps_of(np.float32)
    assert rel(tau[:58], tl[:58]) <= 50 * 60 * eps_of(np.float32)


def _gkl(A):
    """Golub–Kahan–Lanczos bidiagonalisation from v₁ = e₁, full
    reorthogonalisation (m ≥ n): |diagonal| and |superdiagonal|."""
    m, n = A.shape
    U = np.zeros((m, n))
    V = np.zeros((n, n))
    a = np.zeros(n)
    b = np.zeros(max(n - 1, 0))
    V[0, 0] = 1.0
    for k in range(n):
        u = A @ V[:, k] - (b[k - 1] * U[:, k - 1] if k > 0 else 0)
        u -= U[:, :k] @ (U[:, :k].T @ u)
        a[k] = np.linalg.norm(u)
        U[:, k] = u / a[k]
        if k + 1 < n:
            v = A.T @ U[:, k] - a[k] * V[:, k]
            v -= V[:, :k + 1] @ (V[:, :k + 1].T @ v)
            b[k] = np.linalg.norm(v)
            V[:, k + 1] = v / b[k]
    return a, b


@pytest.mark.parametrize("m,n", [(1, 1), (2, 2), (9, 6), (6, 9), (50, 50), (120, 40), (40, 120)])
def test_bidiag_invariants(m, n, rng):
    A = rng.standard_normal((m, n))
    F, tl, tr = O.bidiag(A)
    Ql, Qr, B = O.bidiag_q(F, tl, tr)
    k = max(m, n)
    tol = 10 * k * eps_of(np.float64)
    assert np.linalg.norm(Ql.T @ Ql - np.eye(m)) <= tol
    assert np.linalg.norm(Qr.T @ Qr - np.eye(n)) <= tol
    assert rel(Ql.T @ A @ Qr, B) <= tol
    assert rel(np.linalg.svd(B, compute_uv=False), np.linalg.svd(A, compute_uv=False)) <= tol


@pytest.mark.parametrize("m,n", [(9, 6), (60, 60), (100, 30)])
def test_bidiag_against_lanczos_and_qr(m, n, rng):
    A = rng.standard_normal((m, n))
    F, tl, tr = O.bidiag(A)
    a, b = _gkl(A)
    assert rel(np.abs(np.diag(F)[:n]), a) <= 1e-10
    assert rel(np.abs(np.diag(F, 1)[:n - 1]), b) <= 1e-10
    qr, tq, _, info = lp.dgeqrf(A)
    assert info == 0
    assert abs(tl[0] - tq[0]) <= 10 * eps_of(np.float64)
    assert rel(F[1:, 0], qr[1:, 0]) <= 10 * m * eps_of(np.float64)
    assert F[0, 0] == pytest.approx(qr[0, 0], rel=1e-14)


def test_bidiag_wide_is_transpose_problem(rng):
    """For m < n the lower bidiagonal of A has the singular values of the
    upper bidiagonal of Aᵀ (the reference's wide branch mirrors the tall
    one), and B = Qlᵀ A Qr."""
    A = rng.standard_normal((30, 70))
    F, tl, tr = O.bidiag(A)
    Ft, tlt, trt = O.bidiag(A.T.copy())
    assert rel(np.abs(np.diag(F)), np.abs(np.diag(Ft))) <= 1e-12
    assert rel(np.abs(np.diag(F, -1)), np.abs(np.diag(Ft, 1))) <= 1e-12


def test_planted_slip_fails(rng):
    """The rank-2 update without ξ x xᵀ is a different, wrong reduction."""
    n = 12
    A = sym(n, rng)
    AS, tau = O.symtri(A, "L")
    Q = O.symtri_q(AS, tau, "L")
    # redo step 1 by hand without the ξ term: the trailing block differs
    x = AS[1:, 0].copy()
    x[0] = 1.0
    At = A[1:, 1:]
    u = tau[0] * At @ x
    bad = At - np.outer(x, u) - np.outer(u, x)
    good = At - np.outer(x, u) - np.outer(u, x) + tau[0] * (x @ u) * np.outer(x, x)
    H = np.eye(n - 1) - tau[0] * np.outer(x, x)
    assert rel(H @ At @ H, good) <= 1e-13
    assert rel(H @ At @ H, bad) > 1e-3
    assert rel(Q.T @ A @ Q, np.diag(np.diag(AS)) + np.diag(np.diag(AS, -1), -1)
               + np.diag(np.diag(AS, -1), 1)) <= 1e-13


@pytest.mark.parametrize("n", [20, 64])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symtri_quad_is_the_truth_both_fp64_forms_approach(n, uplo, rng):
    """The __float128 evaluation (the BigFloat stand-in) is what the FP64
    oracle and LAPACK dsytrd both approximate: each within 1e-11 of it,
    and it satisfies Qᵀ A Q = T to float64 rounding."""
    A = sym(n, rng)
    tru, ttru = O.symtri_quad(A, uplo)
    ref, tref = O.symtri(A, uplo)
    c, dl, el, tl, info = lp.dsytrd(A, lower=1 if uplo == "L" else 0)
    tri = np.tril if uplo == "L" else np.triu
    assert rel(tri(ref), tri(tru)) <= 1e-11
    assert rel(tri(c), tri(tru)) <= 1e-11
    Q = O.symtri_q(tru, ttru, uplo)
    d, e = O.symtri_diagonals(tru, uplo)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.linalg.norm(Q.T @ A @ Q - T) <= 10 * n * eps_of(np.float64) * np.linalg.norm(A)


@pytest.mark.parametrize("m,n", [(30, 20), (20, 30)])
def test_bidiag_quad_agrees(m, n, rng):
    A = rng.standard_normal((m, n))
    tru, tl, tr = O.bidiag_quad(A)
    ref, rl, rr = O.bidiag(A)
    assert rel(ref, tru) <= 1e-11 and rel(rl, tl) <= 1e-11 and rel(rr, tr) <= 1e-11

