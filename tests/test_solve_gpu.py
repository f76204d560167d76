"""Solves with the factors on the device: ldiv! with a Cholesky factor
(potrs) and least squares through a Householder QR.  Checked against LAPACK
on the host (numpy.linalg.solve / lstsq) and by what the solutions must
satisfy: the normwise backward error ‖AX − B‖/(‖A‖‖X‖ + ‖B‖) ≤ c·n·eps, and
for least squares the normal equations Aᵀ(AX − B) ≈ 0.  The matrices are
well conditioned (κ ≲ 10), so X itself is compared at c·n·eps too."""
import numpy as np
import pytest

from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def _tdt(dtype):
    return torch.float64 if dtype == np.float64 else torch.float32


def _spd(n, rng):
    B = rng.standard_normal((n, n))
    return B.T @ B / n + np.eye(n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 5, 128, 129, 300, 1000])
@pytest.mark.parametrize("nrhs", [1, 3, 17])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_chol_solve_vs_lapack(dtype, n, nrhs, uplo):
    gla = _gla()
    rng = np.random.default_rng(n * 100 + nrhs + (1 if uplo == "U" else 0))
    A = _spd(n, rng).astype(dtype)
    B = rng.standard_normal((n, nrhs)).astype(dtype)
    Ag = gla.colmajor(A, _tdt(dtype))
    gla.cholRecursive_(Ag, getattr(gla.Val, uplo))
    # the triangle not holding the factor keeps A's entries: it must not be read
    Bg = gla.colmajor(B, _tdt(dtype))
    gla.chol_solve_(Ag, Bg, getattr(gla.Val, uplo))
    X = gla.to_numpy(Bg).astype(np.float64)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    eps = eps_of(dtype)
    be = np.linalg.norm(A64 @ X - B64) / (np.linalg.norm(A64) * np.linalg.norm(X) +
                                          np.linalg.norm(B64))
    assert be <= 10 * n * eps, be
    assert rel(X, np.linalg.solve(A64, B64)) <= 50 * n * eps


def test_chol_solve_vector_and_wrapped_factor():
    gla = _gla()
    rng = np.random.default_rng(8)
    n = 77
    A = _spd(n, rng)
    b = rng.standard_normal(n)
    L = gla.cholRecursive_(gla.colmajor(A), gla.Val.L)          # LowerTriangular wrapper
    x = gla.chol_solve_(L, torch.from_numpy(b).cuda())
    assert x.shape == (n,)
    assert rel(x.cpu().numpy(), np.linalg.solve(A, b)) <= 50 * n * eps_of(np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n", [(1, 1), (10, 5), (129, 129), (300, 128), (1000, 257),
                                 (2000, 600)])
@pytest.mark.parametrize("nrhs", [1, 4])
def test_qr_least_squares_vs_lapack(dtype, m, n, nrhs):
    gla = _gla()
    rng = np.random.default_rng(m + 7 * n + nrhs)
    A = rng.standard_normal((m, n)).astype(dtype)
    B = rng.standard_normal((m, nrhs)).astype(dtype)
    F = gla.qrBlocked_(gla.colmajor(A, _tdt(dtype)), 64)
    Bg = gla.colmajor(B, _tdt(dtype))
    X = gla.to_numpy(gla.qr_solve_(F, Bg)).astype(np.float64)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    eps = eps_of(dtype)
    Xr = np.linalg.lstsq(A64, B64, rcond=None)[0]
    assert rel(X, Xr) <= 50 * m * eps, rel(X, Xr)
    # normal equations: Aᵀ r ≈ 0 relative to ‖A‖ (‖A‖‖X‖ + ‖B‖)
    r = A64 @ X - B64
    scale = np.linalg.norm(A64) * (np.linalg.norm(A64) * np.linalg.norm(X) + np.linalg.norm(B64))
    assert np.linalg.norm(A64.T @ r) <= 10 * m * eps * scale
    if m == n:
        assert np.linalg.norm(r) <= 10 * m * eps * (np.linalg.norm(A64) * np.linalg.norm(X) +
                                                    np.linalg.norm(B64))


def test_qr_solve_singular_r_raises():
    gla = _gla()
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 6))
    A[:, 2] = 0.0                      # reflector of a zero column: τ = 0, R[2,2] = 0
    F = gla.qrBlocked_(gla.colmajor(A), 4)
    with pytest.raises(gla.ArgumentError, match="singular"):
        gla.qr_solve_(F, gla.colmajor(rng.standard_normal((40, 1))))


def test_solve_argument_errors():
    gla = _gla()
    F = gla.colmajor(np.eye(4))
    with pytest.raises(gla.DimensionMismatch):
        gla.chol_solve_(F, gla.colmajor(np.ones((5, 1))))
    q = gla.qrBlocked_(gla.colmajor(np.ones((3, 5))), 2)
    with pytest.raises(gla.ArgumentError):
        gla.qr_solve_(q, gla.colmajor(np.ones((3, 1))))
