"""The two-sided Householder reductions on the GPU through the C ABI:

- symtri! (symtriLower!/symtriUpper!, src/eigenSelfAdjoint.jl:450-564)
  against the oracle: the reduced triangle (tridiagonal and reflectors) and
  τ measured against the reference evaluated in __float128 (the truth):
  within 10·n·eps of it, or no further from it than twice the FP64
  oracle's own distance (n ≤ 300; above, the FP64 oracle at 10·n·eps)
  (the reduction's outputs are only as well
  conditioned as the input makes them: for some inputs the FP64 oracle and
  LAPACK dsytrd are themselves 2e-13 from the truth at n = 64, beyond
  10·n·eps); backward: Qᵀ A Q = T to 10·n·eps; the other triangle untouched; the reference's own test
  (test/eigenselfadjoint.jl:61-67: Q orthogonal, Qᵀ A Q = T) with Q formed
  from the device's reflectors; at n = 3000 the tridiagonal's eigenvalues
  against LAPACK's of A.
- bidiagonalize! (src/svd.jl:328-381), tall and wide, against the oracle
  (B, reflectors, τl, τr to 10·max(m,n)·eps) and through Qlᵀ A Qr = B and
  the singular values (test/svd.jl:100-112).
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def sym(n, rng, dtype=np.float64):
    G = rng.standard_normal((n, n))
    return (G + G.T).astype(dtype)


def _symtri(A, uplo):
    gla = _gla()
    X = gla.colmajor(A)
    F = gla.symtri_(gla.Hermitian(X, uplo))
    torch.cuda.synchronize()
    return gla.to_numpy(F.factors), F.tau.cpu().numpy(), F


SIZES = [1, 2, 3, 4, 10, 31, 32, 33, 34, 35, 64, 65, 66, 100, 129, 300, 700]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symtri_vs_oracle(n, uplo):
    rng = np.random.default_rng(500 + n)
    A = sym(n, rng)
    got, tau, F = _symtri(A, uplo)
    ref, tref = O.symtri(A, uplo)
    tri = np.tril if uplo == "L" else np.triu
    tol = 10 * n * eps_of(np.float64)
    if n <= 300:
        tru, ttru = O.symtri_quad(A, uplo)
        assert rel(tri(got), tri(tru)) <= max(tol, 2 * rel(tri(ref), tri(tru)))
        if n > 1:
            assert rel(tau, ttru) <= max(tol, 2 * rel(tref, ttru))
    else:   # the quad evaluation takes minutes here: the FP64 oracle at 10·n·eps
        assert rel(tri(got), tri(ref)) <= tol
        assert rel(tau, tref) <= tol
    if n > 2:
        Q = O.symtri_q(got, tau, uplo)
        d, e = (x.cpu().numpy() for x in F.diagonals)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.linalg.norm(Q.T @ A @ Q - T) <= tol * np.linalg.norm(A)
    other = np.triu if uplo == "L" else np.tril
    k = 1 if uplo == "L" else -1
    assert np.array_equal(other(got, k), other(A, k)), "other triangle changed"


@pytest.mark.parametrize("n", [50, 257])
def test_symtri_float32(n):
    """FP32 against the FP64 oracle of the same input (the truth at FP32's
    scale): within 10·n·eps of FP32, or of twice the FP32 oracle's own
    error — two FP32 evaluations that sum in different orders can each sit
    near that bound, so the FP32 oracle is not the yardstick itself."""
    rng = np.random.default_rng(n)
    A = sym(n, rng, np.float32)
    tol = 10 * n * eps_of(np.float32)
    for uplo in ("L", "U"):
        got, tau, _ = _symtri(A, uplo)
        ref, tref = O.symtri(A, uplo)
        tru, ttru = O.symtri(A.astype(np.float64), uplo)
        tri = np.tril if uplo == "L" else np.triu
        assert rel(tri(got), tri(tru)) <= max(tol, 2 * rel(tri(ref), tri(tru)))
        assert rel(tau, ttru) <= max(tol, 2 * rel(tref, ttru))


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symtri_reference_test(uplo):
    """test/eigenselfadjoint.jl:61-67 (real entries; A = V Λ Vᵀ with Λ
    log-spaced over 1e-8..1): Q'Q ≈ I ≈ QQ', Q' A Q ≈ T."""
    n = 120
    rng = np.random.default_rng(61)
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = 10.0 ** np.linspace(-8, 0, n)
    A = V @ np.diag(lam) @ V.T
    A = (A + A.T) / 2
    got, tau, F = _symtri(A, uplo)
    Q = O.symtri_q(got, tau, uplo)
    d, e = (x.cpu().numpy() for x in F.diagonals)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    tol = 10 * n * eps_of(np.float64)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= tol
    assert np.linalg.norm(Q @ Q.T - np.eye(n)) <= tol
    assert rel(Q.T @ A @ Q, T) <= tol


def test_symtri_large_eigenvalues():
    """n = 3000: T's eigenvalues equal A's (LAPACK dsyevd) to 10·n·eps·‖A‖."""
    n = 3000
    A = sym(n, np.random.default_rng(3000))
    got, tau, F = _symtri(A, "L")
    d, e = (x.cpu().numpy() for x in F.diagonals)
    lt = sl.eigvalsh_tridiagonal(d, e)
    la = np.linalg.eigvalsh(A)
    assert np.max(np.abs(lt - la)) <= 10 * n * eps_of(np.float64) * np.max(np.abs(la))


def test_symtri_argument_errors():
    gla = _gla()
    with pytest.raises(gla.DimensionMismatch):
        gla.symtri_(gla.Hermitian(gla.colmajor(np.zeros((3, 4))), "L"))
    with pytest.raises(TypeError):
        gla.symtri_(gla.colmajor(np.eye(3)))


# ---------------------------------------------------------- bidiagonalize! --
def _bidiag(A):
    gla = _gla()
    X = gla.colmajor(A)
    F = gla.bidiagonalize_(X)
    torch.cuda.synchronize()
    return gla.to_numpy(F.factors), F.taul.cpu().numpy(), F.taur.cpu().numpy(), F


SHAPES = [(1, 1), (2, 2), (1, 5), (5, 1), (3, 2), (2, 3), (10, 10), (40, 33), (33, 40), (64, 64),
          (65, 64), (64, 65), (100, 70), (70, 100), (129, 129), (300, 128), (128, 300),
          (500, 500)]


@pytest.mark.parametrize("m,n", SHAPES)
def test_bidiag_vs_oracle(m, n):
    """Measured against the reference evaluated in __float128, as for
    symtri!: within 10·max(m,n)·eps, or no further from it than twice the
    FP64 oracle (max(m,n) ≤ 500; above, the FP64 oracle at 10·max(m,n)·eps);
    backward: Qlᵀ A Qr = B to 10·max(m,n)·eps."""
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    got, tl, tr, F = _bidiag(A)
    ref, rl, rr = O.bidiag(A)
    k = max(m, n)
    tol = 10 * k * eps_of(np.float64)
    if k <= 500:
        tru, ql, qr = O.bidiag_quad(A)
        assert rel(got, tru) <= max(tol, 2 * rel(ref, tru))
        if len(ql):
            assert rel(tl, ql) <= max(tol, 2 * rel(rl, ql))
        if len(qr):
            assert rel(tr, qr) <= max(tol, 2 * rel(rr, qr))
    else:
        assert rel(got, ref) <= tol
        assert rel(tl, rl) <= tol and rel(tr, rr) <= tol
    Ql, Qr, B = O.bidiag_q(got, tl, tr)
    assert np.linalg.norm(Ql.T @ A @ Qr - B) <= tol * np.linalg.norm(A)
    dv, ev = (x.cpu().numpy() for x in F.bidiagonal)
    assert np.array_equal(dv, np.diag(got)[:min(m, n)])


@pytest.mark.parametrize("m,n", [(200, 150), (150, 200)])
def test_bidiag_float32(m, n):
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)).astype(np.float32)
    got, tl, tr, _ = _bidiag(A)
    ref, rl, rr = O.bidiag(A)
    tru, tlt, trt = O.bidiag(A.astype(np.float64))
    tol = 10 * max(m, n) * eps_of(np.float32)
    # against the FP64 oracle, within the bar or twice the FP32 oracle's error
    assert rel(got, tru) <= max(tol, 2 * rel(ref, tru))
    assert rel(tl, tlt) <= max(tol, 2 * rel(rl, tlt))
    assert rel(tr, trt) <= max(tol, 2 * rel(rr, trt))


def test_bidiag_reference_svdvals():
    """test/svd.jl:100-112: the bidiagonal's singular values are A's (the
    8×8 matrix of that test), and a larger 2000×1500 case against LAPACK."""
    A = np.array([
        [0.3, 0.0, 0.0, 0.0, 0.0, 0.2, 0.3, 0.0],
        [0.0, 0.0, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0],
        [0.0, -0.2, 0.0, 0.0, 0.0, 0.0, 0.0, -0.2],
        [0.3, 0.0, 0.0, 0.0, 0.0, 0.2, 0.4, 0.0],
        [0.0, 0.4, -0.2, 0.0, 0.0, 0.0, 0.0, 0.3],
        [0.2, 0.0, 0.0, 0.0, 0.0, 0.0, 0.2, 0.0],
        [0.0, 0.0, 0.0, 0.1, 0.0, 0.0, 0.0, 0.0],
        [0.0, 0.3, -0.2, 0.0, 0.0, 0.0, 0.0, 0.3],
    ])
    for M in (A, np.random.default_rng(2).standard_normal((2000, 1500))):
        got, tl, tr, F = _bidiag(M)
        dv, ev = (x.cpu().numpy() for x in F.bidiagonal)
        k = min(M.shape)
        B = np.diag(dv) + np.diag(ev, 1)
        s = np.linalg.svd(B, compute_uv=False)
        sref = np.linalg.svd(M, compute_uv=False)
        assert np.max(np.abs(s - sref)) <= 10 * max(M.shape) * eps_of(np.float64) * sref[0]
