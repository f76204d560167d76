"""Pins the oracle's LDLᵀ (src/ldlt.jl) on the CPU, independently of its own
formulas:

- the docstring examples of ldlt (src/ldlt.jl:170-200): [1 1; 1 −1] gives
  L = [1 0; 1 1], D = diag(1, −2); [1 1; 1 1] gives D = diag(1, 0);
- D's entries are the ratios of consecutive leading principal minors
  (exact rational arithmetic on integer matrices, any symmetric indefinite
  matrix with nonzero minors);
- for SPD matrices, L = L_c diag(L_c)⁻¹ and D = diag(L_c)² with L_c from
  LAPACK dpotrf;
- the reference's own tests (test/ldlt.jl): L D Lᵀ ≈ B for B = A Aᵀ with
  A n×(n−2), and (B + I)·(F \\ o) ≈ o, for n ∈ {5, 50, 500} and both uplo;
- the result does not depend on the block size (unblocked :16-62 against
  blocked :80-148), and uplo 'U' is the transpose of uplo 'L' to rounding.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from conftest import eps_of, rel


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_docstring_examples(uplo):
    F = O.ldlt(np.array([[1.0, 1.0], [1.0, -1.0]]), uplo)
    L, d = O.ldlt_factors(F, uplo)
    assert np.array_equal(L, [[1.0, 0.0], [1.0, 1.0]])
    assert np.array_equal(d, [1.0, -2.0])
    F = O.ldlt(np.array([[1.0, 1.0], [1.0, 1.0]]), uplo)
    L, d = O.ldlt_factors(F, uplo)
    assert np.array_equal(L, [[1.0, 0.0], [1.0, 1.0]])
    assert np.array_equal(d, [1.0, 0.0])


def _det(M):
    """Exact determinant of a small integer matrix (Fraction elimination)."""
    M = [[Fraction(int(x)) for x in row] for row in M]
    n = len(M)
    det = Fraction(1)
    for k in range(n):
        p = next((r for r in range(k, n) if M[r][k] != 0), None)
        if p is None:
            return Fraction(0)
        if p != k:
            M[k], M[p] = M[p], M[k]
            det = -det
        det *= M[k][k]
        for r in range(k + 1, n):
            f = M[r][k] / M[k][k]
            for c in range(k, n):
                M[r][c] -= f * M[k][c]
    return det


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_d_is_ratio_of_leading_minors(seed, uplo):
    rng = np.random.default_rng(seed)
    n = 9
    while True:
        S = rng.integers(-6, 7, size=(n, n))
        S = S + S.T
        minors = [Fraction(1)] + [_det(S[:k, :k]) for k in range(1, n + 1)]
        if all(m != 0 for m in minors[1:]):
            break
    assert any(minors[k] / minors[k - 1] < 0 for k in range(1, n + 1)), "indefinite"
    d_exact = np.array([float(minors[k] / minors[k - 1]) for k in range(1, n + 1)])
    for bs in (1, 4, 16):
        F = O.ldlt(S.astype(np.float64), uplo, bs)
        _, d = O.ldlt_factors(F, uplo)
        assert rel(d, d_exact) <= 1e-11, (bs, d, d_exact)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spd_matches_lapack_cholesky(dtype, rng):
    n = 300
    B = rng.standard_normal((n, n))
    A = (B.T @ B + n * np.eye(n)).astype(dtype)
    Lc = sl.cholesky(A.astype(np.float64), lower=True)
    for uplo in ("L", "U"):
        F = O.ldlt(A, uplo)
        L, d = O.ldlt_factors(F, uplo)
        tol = 10 * n * eps_of(dtype)
        assert rel(d, np.diag(Lc) ** 2) <= tol
        assert rel(L, Lc / np.diag(Lc)) <= tol


@pytest.mark.parametrize("n", [5, 50, 500])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_reference_tests_reconstruction_and_solve(n, uplo):
    """test/ldlt.jl:5-20 with real entries: B = A Aᵀ, A n×(n−2) (singular
    PSD, so the last pivots are rounding noise)."""
    rng = np.random.default_rng(123 + n)
    A = rng.standard_normal((n, n - 2))
    B = A @ A.T
    F = O.ldlt(B, uplo)
    L, d = O.ldlt_factors(F, uplo)
    assert rel(L @ np.diag(d) @ L.T, B) <= 1e-10
    Bfr = B + np.eye(n)
    F = O.ldlt(Bfr, uplo)
    o = np.ones(n)
    x = O.ldlt_solve(F, o, uplo)
    assert rel(Bfr @ x, o) <= 1e-10
    assert rel(x, np.linalg.solve(Bfr, o)) <= 1e-8


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_blocksize_invariance_and_triangle_untouched(uplo, rng):
    n = 70
    S = rng.standard_normal((n, n))
    S = S + S.T + 12 * np.diag(np.sign(rng.standard_normal(n)))   # indefinite, modest growth
    ref = O.ldlt_unblocked(S, uplo)
    tri = np.tril if uplo == "L" else np.triu
    other = np.triu if uplo == "L" else np.tril
    for bs in (1, 5, 16, 33, 69, 70, 200):
        F = O.ldlt(S, uplo, bs)
        assert rel(tri(F), tri(ref)) <= 100 * n * eps_of(np.float64), bs
        assert np.array_equal(other(F, 1 if uplo == "L" else -1),
                              other(S, 1 if uplo == "L" else -1))


def test_upper_is_transpose_of_lower(rng):
    n = 120
    S = rng.standard_normal((n, n))
    S = S + S.T + 60 * np.diag(np.sign(rng.standard_normal(n)))   # indefinite, small growth
    Fl = O.ldlt(S, "L")
    Fu = O.ldlt(S, "U")
    assert rel(np.triu(Fu), np.tril(Fl).T) <= 100 * n * eps_of(np.float64)


def test_planted_slip_fails_the_minor_check():
    """A trailing update without D (A22 −= L21 L21ᵀ, a plausible slip) gives
    a different D, which the leading-minor check detects."""
    S = np.array([[4.0, 2, 2], [2, -3, 1], [2, 1, 5]])
    d_exact = [4.0, -3 - 1.0, None]
    F = O.ldlt(S, "L", 1)
    _, d = O.ldlt_factors(F, "L")
    assert d[0] == d_exact[0] and d[1] == d_exact[1]
    # the slip: Schur complement formed with l l' instead of l δ l'
    l = S[1:, 0] / S[0, 0]
    slip = S[1:, 1:] - np.outer(l, l)
    assert slip[0, 0] != d[1]
