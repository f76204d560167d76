"""Complex (ComplexF64/ComplexF32) Cholesky on the GPU through the C ABI:
cholUnblocked!, cholBlocked!, cholRecursive! (src/cholesky.jl:3-55, her!/
herk! updates per src/juliaBLAS.jl:41-48,78-86) against the oracle's
complex restatement and LAPACK zpotrf at 10·n·eps; the reference's own
complex test (test/cholesky.jl:11-27: A = complex(rand, rand), A'A,
blocksizes 5 and 10, cutoffs 1 and 4); rankUpdate! for complex C; the
opposite triangle untouched; the complex square root of a negative pivot
raising no error, as in the reference."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CTYPES = [np.complex128, np.complex64]


def _gla():
    import gla

    return gla


def hpd(n, rng, dtype=np.complex128):
    """BᴴB + n·I, made exactly Hermitian (real diagonal)."""
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = B.conj().T @ B + n * np.eye(n)
    return (0.5 * (A + A.conj().T)).astype(dtype)


def _run(A, variant, uplo="L", param=None):
    gla = _gla()
    X = gla.colmajor(A)
    if variant == "unblocked":
        gla.cholUnblocked_(X, uplo)
    elif variant == "blocked":
        gla.cholBlocked_(X, uplo, param)
    else:
        gla.cholRecursive_(X, uplo, param if param is not None else 1)
    torch.cuda.synchronize()
    return gla.to_numpy(X)


@pytest.mark.parametrize("dtype", CTYPES)
@pytest.mark.parametrize("variant,param", [("unblocked", None), ("blocked", 5), ("blocked", 10),
                                           ("recursive", 1), ("recursive", 4)])
def test_reference_complex_cholesky(dtype, variant, param):
    """test/cholesky.jl:11-27: LowerTriangular(chol*(copy(A'A))) ≈
    cholesky(Hermitian(A'A)).L, here at 10·n·eps against the oracle and
    LAPACK zpotrf."""
    n = 50
    rng = np.random.default_rng(123)
    A = rng.random((n, n)) + 1j * rng.random((n, n))
    AcA = (A.conj().T @ A).astype(dtype)
    got = np.tril(_run(AcA, variant, "L", param))
    ref, info = O.zchol(AcA, variant, param if param is not None else 1)
    assert info == 0
    tol = 10 * n * eps_of(dtype)
    assert rel(got, np.tril(ref)) <= tol
    assert rel(got, sl.cholesky(AcA.astype(np.complex128), lower=True)) <= tol


@pytest.mark.parametrize("n", [1, 2, 17, 63, 64, 65, 130, 300, 1000])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("dtype", CTYPES)
def test_complex_cholesky_vs_oracle(n, uplo, dtype):
    rng = np.random.default_rng(2000 + n)
    A = hpd(n, rng, dtype)
    got = _run(A, "recursive", uplo)
    ref, info = O.zchol(A, "recursive", 1)
    assert info == 0
    tol = 10 * n * eps_of(dtype)
    if uplo == "L":
        assert rel(np.tril(got), np.tril(ref)) <= tol
        assert np.array_equal(np.triu(got, 1), np.triu(A, 1))
    else:
        assert rel(np.triu(got), np.tril(ref).conj().T) <= tol
        assert np.array_equal(np.tril(got, -1), np.tril(A, -1))
    d = np.diag(got)
    assert np.all(d.imag == 0) and np.all(d.real > 0)


@pytest.mark.parametrize("bs", [1, 3, 16, 64, 100])
def test_complex_blocked_blocksizes(bs):
    n = 200
    A = hpd(n, np.random.default_rng(bs))
    got = np.tril(_run(A, "blocked", "L", bs))
    ref, _ = O.zchol(A, "blocked", bs)
    assert rel(got, np.tril(ref)) <= 10 * n * eps_of(np.complex128)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("dtype", CTYPES)
def test_complex_rank_update(uplo, dtype):
    """rankUpdate!(Hermitian(C, uplo), A, -1) for complex: the uplo triangle
    against the oracle's herk, the other untouched, the diagonal real."""
    gla = _gla()
    rng = np.random.default_rng(7)
    n, k = 150, 37
    Cm = hpd(n, rng, dtype)
    A = (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))).astype(dtype)
    Cg = gla.colmajor(Cm)
    gla.rankUpdate_(gla.Hermitian(Cg, uplo), gla.colmajor(A), -1.0)
    got = gla.to_numpy(Cg)
    if uplo == "L":
        ref = O.zherk_l(Cm, A, -1.0)
        assert rel(np.tril(got), np.tril(ref)) <= 10 * n * eps_of(dtype)
        assert np.array_equal(np.triu(got, 1), np.triu(Cm, 1))
    else:
        full = Cm - A @ A.conj().T
        assert rel(np.triu(got), np.triu(full)) <= 10 * n * eps_of(dtype)
        assert np.array_equal(np.tril(got, -1), np.tril(Cm, -1))
    assert np.all(np.diag(got).imag == 0)


def test_complex_negative_pivot_is_not_an_error():
    """The complex sqrt of a negative pivot (src/cholesky.jl:5) returns an
    imaginary root: no exception; the first such pivot is reported."""
    gla = _gla()
    A = hpd(40, np.random.default_rng(9))
    A[20, 20] = -1e6
    X = gla.colmajor(A)
    gla.cholRecursive_(X, gla.Val.L)          # no DomainError
    torch.cuda.synchronize()
    assert gla.last_pivot_info() == 21
    ref, info = O.zchol(A, "unblocked")
    assert info == 21
    got = gla.to_numpy(X)
    assert rel(np.tril(got)[:20, :20], np.tril(ref)[:20, :20]) <= 10 * 40 * eps_of(np.complex128)
    assert got[20, 20].real == 0 and got[20, 20].imag > 0


def test_complex_cholesky_4096_residual():
    """n = 4096 ComplexF64: ‖A − L Lᴴ‖/‖A‖ on sampled rows ≤ 10·n·eps."""
    n = 4096
    g = torch.Generator(device="cuda").manual_seed(4096)
    B = torch.complex(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    A = B.conj().t() @ B
    A.diagonal().add_(n)
    A = 0.5 * (A + A.conj().t())
    gla = _gla()
    W = A.t().contiguous().t()
    gla.cholRecursive_(W, gla.Val.L)
    L = torch.tril(W)
    rows = torch.from_numpy(np.sort(np.random.default_rng(1).choice(n, 64, replace=False))).cuda()
    R = A[rows] - L[rows] @ L.conj().t()
    err = float(torch.linalg.norm(R) / torch.linalg.norm(A[rows]))
    assert err <= 10 * n * eps_of(np.float64), err


# ----------------------------------------------------------------- QR -------
QR_SHAPES = [(10, 5), (10, 10), (5, 10), (100, 50), (100, 100), (50, 100)]
QR_BS = [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33]


def _cmat(m, n, rng, dtype):
    return (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(dtype)


@pytest.mark.parametrize("m,n", QR_SHAPES)
@pytest.mark.parametrize("dtype", CTYPES)
def test_complex_qr_reference_grid(m, n, dtype):
    """test/qr.jl:7-29 in complex: for each block size, the factors equal
    qrUnblocked!'s (the oracle, which LAPACK zgeqrf pins), (Qᴴ A) = R and
    Qᴴ(Q A) = A through the device's compact-WY block."""
    gla = _gla()
    rng = np.random.default_rng(m * 100 + n)
    A = _cmat(m, n, rng, dtype)
    ref, tref = O.zqr_unblocked(A)
    tol = 10 * max(m, n) * eps_of(dtype)
    F = gla.qrUnblocked_(gla.colmajor(A))
    assert rel(gla.to_numpy(F.factors), ref) <= tol
    assert rel(F.tau.cpu().numpy(), tref) <= tol
    for bs in QR_BS:
        F = gla.qrBlocked_(gla.colmajor(A), bs)
        got = gla.to_numpy(F.factors)
        assert rel(got, ref) <= tol, bs
        assert rel(F.tau.cpu().numpy(), tref) <= tol, bs
        H = F[gla.QBlocked]
        Ad = gla.colmajor(A)
        QhA = gla.to_numpy(gla.adjoint(H) @ Ad)
        k = min(m, n)
        assert rel(QhA[:k], np.triu(got)[:k]) <= tol, bs
        back = gla.to_numpy(gla.adjoint(H) @ (H @ Ad))
        assert rel(back, A) <= tol, bs


@pytest.mark.parametrize("dtype", CTYPES)
def test_complex_qr_larger(dtype):
    gla = _gla()
    m, n = 1000, 600
    A = _cmat(m, n, np.random.default_rng(5), dtype)
    ref, tref = O.zqr_unblocked(A)
    F = gla.qrBlocked_(gla.colmajor(A), 32)
    tol = 10 * m * eps_of(dtype)
    assert rel(gla.to_numpy(F.factors), ref) <= tol
    assert rel(F.tau.cpu().numpy(), tref) <= tol
    T = gla.to_numpy(F[gla.QBlocked].T.data)
    V = np.tril(ref[:, :n], -1) + np.eye(m, n)
    # T against the explicit reflector product on a few columns of I
    E = np.eye(m, 7, dtype=np.complex128)
    Qe = E.copy()
    for j in range(n - 1, -1, -1):
        v = V[:, j].astype(np.complex128)
        Qe = Qe - tref[j] * np.outer(v, v.conj() @ Qe)
    assert rel(E - V @ (T @ (V.conj().T @ E)), Qe) <= tol
