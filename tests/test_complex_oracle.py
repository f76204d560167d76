"""Pins the oracle's complex restatements on the CPU:

- cholUnblocked!/cholBlocked!/cholRecursive! with her!/herk! against LAPACK
  zpotrf/cpotrf (same factor: lower, real positive diagonal), on the
  reference's own complex test matrix (test/cholesky.jl:11-27) and on
  BᴴB + n·I; the closed form [[4, 2i], [−2i, 5]] -> L = [[2, 0], [−i, 2]];
- the complex square root of a negative pivot (no error, imaginary root);
- herk against numpy's A Aᴴ;
- the complex reflector! and qrUnblocked! against LAPACK zgeqrf (zlarfg uses
  the same convention: β = −copysign(‖x‖, real(x₁)), τ = (β − x₁)/β), and
  Qᴴ A = R with Q formed from the reflectors.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.linalg.lapack as lp

import oracle as O
from conftest import eps_of, rel


def hpd(n, rng, dtype=np.complex128):
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = B.conj().T @ B + n * np.eye(n)
    return (0.5 * (A + A.conj().T)).astype(dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("variant,param", [("unblocked", 0), ("blocked", 5), ("blocked", 10),
                                           ("recursive", 1), ("recursive", 4)])
def test_zchol_matches_lapack(dtype, variant, param):
    rng = np.random.default_rng(123)
    n = 50
    A = rng.random((n, n)) + 1j * rng.random((n, n))
    AcA = (A.conj().T @ A).astype(dtype)
    F, info = O.zchol(AcA, variant, param)
    assert info == 0
    Lref = sl.cholesky(AcA.astype(np.complex128), lower=True)
    assert rel(np.tril(F), Lref) <= 10 * n * eps_of(dtype)
    B = hpd(120, rng, dtype)
    F, info = O.zchol(B, variant, param)
    assert rel(np.tril(F), sl.cholesky(B.astype(np.complex128), lower=True)) <= 10 * 120 * eps_of(dtype)
    assert np.array_equal(np.triu(F, 1), np.triu(B, 1))


def test_zchol_closed_form():
    A = np.array([[4, 2j], [-2j, 5]], dtype=np.complex128)
    for variant in ("unblocked", "blocked", "recursive"):
        F, info = O.zchol(A, variant, 1)
        assert info == 0
        assert np.allclose(np.tril(F), [[2, 0], [-1j, 2]], rtol=0, atol=1e-15)


def test_zchol_negative_pivot_takes_the_complex_root():
    A = np.array([[-4, 0], [0, 1]], dtype=np.complex128)
    F, info = O.zchol(A, "unblocked")
    assert info == 1
    assert F[0, 0] == 2j


def test_zherk(rng):
    n, k = 40, 9
    C = hpd(n, rng)
    A = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    got = O.zherk_l(C, A, -2.0)
    ref = C - 2.0 * A @ A.conj().T
    assert rel(np.tril(got), np.tril(ref)) <= 10 * n * eps_of(np.complex128)
    assert np.all(np.diag(got).imag == 0)


@pytest.mark.parametrize("m,n", [(1, 1), (6, 4), (30, 30), (40, 25), (25, 40)])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_zqr_matches_lapack(m, n, dtype, rng):
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(dtype)
    F, tau = O.zqr_unblocked(A)
    f = lp.zgeqrf if dtype == np.complex128 else lp.cgeqrf
    qr, tq, _, info = f(A)
    assert info == 0
    tol = 10 * max(m, n) * eps_of(dtype)
    if m > 1:
        assert rel(F, qr) <= tol and rel(tau, tq) <= tol
    k = min(m, n)
    Q = np.eye(m, dtype=np.complex128)
    for j in range(k):
        v = np.zeros(m, dtype=np.complex128)
        v[j] = 1
        v[j + 1:] = F[j + 1:, j]
        Q = Q @ (np.eye(m) - tau[j] * np.outer(v, v.conj()))
    assert rel(Q.conj().T @ A, np.triu(F)) <= tol


def test_complex_reflector_known_answer():
    """x = [3+4i, 12]: ‖x‖ = 13, ν = copysign(13, 3) = 13, β = −13,
    ξ₁ = 16+4i, τ = ξ₁/ν = (16+4i)/13, v₂ = 12/(16+4i)."""
    F, tau = O.zqr_unblocked(np.array([[3 + 4j], [12 + 0j]]))
    assert F[0, 0] == pytest.approx(-13)
    assert tau[0] == pytest.approx((16 + 4j) / 13)
    assert F[1, 0] == pytest.approx(12 / (16 + 4j))


def test_reference_complex_tblock_needs_the_conjugate(rng):
    """src/qr.jl:72 forms T[i,j] = τ_i (F[j,i] + Σ conj(F[k,i]) F[k,j]);
    for complex reflectors the first term must be conj(F[j,i]) for
    I − V T Vᴴ = H₁⋯H_k.  This is why the complex qrBlocked! here builds
    the conjugated T (equal to LAPACK zlarft's) rather than the literal
    one, whose trailing updates would differ from qrUnblocked!."""
    A = rng.standard_normal((12, 5)) + 1j * rng.standard_normal((12, 5))
    F, tau = O.zqr_unblocked(A)
    m, k = 12, 5
    V = np.tril(F, -1) + np.eye(m, k)
    Q = np.eye(m, dtype=complex)
    for j in range(k):
        Q = Q @ (np.eye(m) - tau[j] * np.outer(V[:, j], V[:, j].conj()))

    def tblock(conj_first):
        T = np.zeros((k, k), dtype=complex)
        for j in range(k):
            for i in range(j):
                first = np.conj(F[j, i]) if conj_first else F[j, i]
                T[i, j] = tau[i] * (first + np.vdot(F[j + 1:, i], F[j + 1:, j]))
        Ui = np.linalg.inv(np.triu(T, 1) + np.eye(k))
        return np.triu(Ui, 1) * tau[None, :] + np.diag(tau)

    assert np.linalg.norm(np.eye(m) - V @ tblock(True) @ V.conj().T - Q) <= 1e-14
    assert np.linalg.norm(np.eye(m) - V @ tblock(False) @ V.conj().T - Q) > 1e-2
