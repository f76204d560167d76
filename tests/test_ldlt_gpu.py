"""ldlt! (src/ldlt.jl:150-157) on the GPU through the C ABI against the
oracle's _ldlt_lower_blocked! / _ldlt_upper_blocked! (:80-148) at the
reference's default block size, 10·n·eps in the factor's triangle; the
other triangle keeps its input; ldiv! and lmul! with the factorisation
(:225-234) against the oracle; the reference's own tests (test/ldlt.jl)
and docstring examples."""
import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def indefinite(n, rng, c=3.0):
    """G + Gᵀ + c√n·diag(±1): about half the pivots negative, modest growth."""
    G = rng.standard_normal((n, n))
    return G + G.T + c * np.sqrt(n) * np.diag(np.sign(rng.standard_normal(n)))


def spd(n, rng):
    B = rng.standard_normal((n, n))
    return B.T @ B + n * np.eye(n)


def _factor(A, uplo, dtype, block=512):
    gla = _gla()
    with gla.tuning(ldlt_block=block):
        X = gla.colmajor(A.astype(dtype))
        F = gla.ldlt_(gla.Hermitian(X, uplo))
        torch.cuda.synchronize()
    return F, gla.to_numpy(F.data)


SIZES = [1, 2, 5, 16, 17, 50, 127, 128, 129, 300, 513, 1000, 1500]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ldlt_vs_oracle(n, uplo, dtype):
    rng = np.random.default_rng(1000 + n)
    A = (indefinite(n, rng) if n % 2 else spd(n, rng)).astype(dtype)
    _, got = _factor(A, uplo, dtype)
    ref = O.ldlt(A, uplo)
    tri = np.tril if uplo == "L" else np.triu
    assert rel(tri(got), tri(ref)) <= 10 * max(n, 1) * eps_of(dtype)
    k = -1 if uplo == "L" else 1
    other = np.triu if uplo == "L" else np.tril
    assert np.array_equal(other(got, -k), other(A, -k)), "other triangle changed"


@pytest.mark.parametrize("block", [128, 256, 512, 1024])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_ldlt_step_widths(block, uplo):
    """The step width ("ldlt_block") does not change the factors."""
    n = 2100
    A = indefinite(n, np.random.default_rng(77))
    _, got = _factor(A, uplo, np.float64, block)
    ref = O.ldlt(A, uplo)
    tri = np.tril if uplo == "L" else np.triu
    assert rel(tri(got), tri(ref)) <= 10 * n * eps_of(np.float64)


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_docstring_examples(uplo):
    """src/ldlt.jl:170-200."""
    for A, d in (([[1.0, 1.0], [1.0, -1.0]], [1.0, -2.0]), ([[1.0, 1.0], [1.0, 1.0]], [1.0, 0.0])):
        F, got = _factor(np.array(A), uplo, np.float64)
        assert np.array_equal(F.d.cpu().numpy(), d)
        assert np.array_equal(F.L.cpu().numpy(), [[1.0, 0.0], [1.0, 1.0]])


@pytest.mark.parametrize("n", [5, 50, 500])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_reference_tests(n, uplo):
    """test/ldlt.jl:5-20 with real entries: L·D·L' ≈ B for B = A Aᵀ, A
    n×(n−2) (singular PSD); for B + I: B·(F \\ o) ≈ o and F·(F \\ o) ≈ o."""
    gla = _gla()
    rng = np.random.default_rng(123 + n)
    A = rng.standard_normal((n, n - 2))
    Brr = A @ A.T
    F, _ = _factor(Brr, uplo, np.float64)
    L = F.L
    rec = L @ torch.diag(F.d) @ L.t()
    assert rel(rec.cpu().numpy(), Brr) <= 1e-10
    Bfr = Brr + np.eye(n)
    F, _ = _factor(Bfr, uplo, np.float64)
    o = torch.ones(n, dtype=torch.float64, device="cuda")
    x = F.ldiv_(o.clone())
    assert rel(Bfr @ x.cpu().numpy(), np.ones(n)) <= 1e-10
    y = F @ x.reshape(-1, 1)
    assert rel(y.cpu().numpy().ravel(), np.ones(n)) <= 1e-10
    assert gla.LDLt is type(F)


@pytest.mark.parametrize("n", [1, 7, 129, 700])
@pytest.mark.parametrize("nrhs", [1, 3])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_solve_and_lmul_vs_oracle(n, nrhs, uplo, dtype):
    gla = _gla()
    rng = np.random.default_rng(n * 10 + nrhs)
    A = indefinite(n, rng).astype(dtype)
    Fref = O.ldlt(A, uplo)
    B = rng.standard_normal((n, nrhs)).astype(dtype)
    F = gla.LDLt(gla.colmajor(Fref), uplo)
    X = F.ldiv_(gla.colmajor(B))
    tol = 10 * n * eps_of(dtype)
    xr = O.ldlt_solve(Fref, B, uplo)
    assert rel(gla.to_numpy(X), xr) <= tol * 10
    Lr, dr = O.ldlt_factors(Fref, uplo)
    Y = F.lmul_(gla.colmajor(B))
    yr = Lr.astype(np.float64) @ np.diag(dr.astype(np.float64)) @ Lr.T.astype(np.float64) @ B
    assert rel(gla.to_numpy(Y), yr) <= tol


def test_zero_pivot_propagates_like_the_reference():
    """δ = 0 divides by zero (no error), as _ldlt_lower! does."""
    A = np.array([[0.0, 1.0, 2.0], [1.0, 3.0, 1.0], [2.0, 1.0, 4.0]])
    _, got = _factor(A, "L", np.float64)
    ref = O.ldlt(A, "L")
    g, r = np.tril(got), np.tril(ref)
    assert np.array_equal(np.isnan(g), np.isnan(r)) and np.array_equal(np.isinf(g), np.isinf(r))
    fin = np.isfinite(r)
    assert np.array_equal(g[fin], r[fin])


def test_argument_errors():
    gla = _gla()
    with pytest.raises(gla.DimensionMismatch):
        gla.ldlt_(gla.Hermitian(gla.colmajor(np.zeros((3, 4))), "L"))
    with pytest.raises(TypeError):
        gla.ldlt_(gla.colmajor(np.eye(3)))
    F = gla.ldlt_(gla.Hermitian(gla.colmajor(np.eye(3)), "L"))
    with pytest.raises(gla.DimensionMismatch):
        F.ldiv_(torch.ones(4, dtype=torch.float64, device="cuda"))
