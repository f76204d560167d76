"""Pins the CPU oracle (oracle/gla_oracle.c) before it is trusted as the
checker: LAPACK (scipy, an independent implementation with the same packed
layout and reflector sign convention), closed-form answers, explicit
reflector products, mpmath at 256 bits, brute-force DomainError pivots and
the committed golden fixtures.  CPU only."""
import os

import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg import lapack

import oracle as O
from conftest import eps_of, rel, spd

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------- Cholesky ----
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 2, 5, 50, 129, 300])
def test_chol_variants_match_lapack(rng, dtype, n):
    A = spd(n, rng, dtype)
    Lref, info = (lapack.dpotrf if dtype == np.float64 else lapack.spotrf)(A, lower=1)
    assert info == 0
    Lref = np.tril(Lref)
    for out, inf in (O.chol_unblocked(A), O.chol_blocked(A, 5), O.chol_blocked(A, 10),
                     O.chol_recursive(A, 1), O.chol_recursive(A, 4)):
        assert inf == 0
        assert rel(np.tril(out), Lref) <= 10 * n * eps_of(dtype)
        assert np.array_equal(np.triu(out, 1), np.triu(A, 1))


def test_chol_reference_case_uniform():
    """test/cholesky.jl:8-27 shape: A = rand(50,50); AcA = A'A."""
    rng = np.random.default_rng(123)
    A = rng.random((50, 50))
    AcA = A.T @ A
    L = np.linalg.cholesky(AcA)
    out, info = O.chol_recursive(AcA, 4)
    assert info == 0 and rel(np.tril(out), L) <= 500 * eps_of(np.float64)


def test_chol_min_matrix_exact():
    n = 64
    i = np.arange(1, n + 1)
    A = np.minimum.outer(i, i).astype(np.float64)
    for out, info in (O.chol_unblocked(A), O.chol_blocked(A, 7), O.chol_recursive(A, 1)):
        assert info == 0
        assert np.array_equal(np.tril(out), np.tril(np.ones((n, n))))


def test_chol_pascal_exact():
    from math import comb

    n = 20
    S = np.array([[comb(i + j, i) for j in range(n)] for i in range(n)], dtype=np.float64)
    Lx = np.array([[comb(i, j) if j <= i else 0 for j in range(n)] for i in range(n)], dtype=np.float64)
    out, info = O.chol_recursive(S)
    assert info == 0
    assert rel(np.tril(out), Lx) <= 10 * n * eps_of(np.float64)


def test_chol_2x2_closed_form():
    out, info = O.chol_unblocked(np.array([[4.0, 2.0], [2.0, 3.0]]))
    assert info == 0
    assert out[0, 0] == 2.0 and out[1, 0] == 1.0 and out[1, 1] == np.sqrt(2.0)


@pytest.mark.parametrize("k", [0, 3, 17, 49])
def test_chol_domain_error_pivot_is_first_negative_schur_pivot(rng, k):
    """Brute force: pivot d_k = det(A[:k+1,:k+1]) / det(A[:k,:k]); the first
    negative one is where Julia's sqrt throws."""
    n = 50
    A = spd(n, rng)
    L = np.linalg.cholesky(A)
    A[k, k] -= L[k, k] ** 2 + 1.0
    dets = [1.0] + [np.linalg.det(A[:i, :i]) for i in range(1, n + 1)]
    first = next(i for i in range(1, n + 1) if dets[i] / dets[i - 1] < 0)
    assert first == k + 1
    for _, info in (O.chol_unblocked(A), O.chol_blocked(A, 10), O.chol_recursive(A, 1),
                    O.chol_recursive(A, 8)):
        assert info == first


def test_chol_quad_matches_mpmath():
    """__float128 restatement (the BigFloat stand-in) against mpmath at 256 bits."""
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 256
    rng = np.random.default_rng(5)
    n = 12
    A = spd(n, rng)
    Lq, info = O.chol_recursive_q(A)
    assert info == 0
    M = mp.matrix([[mp.mpf(float(A[i, j])) for j in range(n)] for i in range(n)])
    Lm = mp.cholesky(M)
    Lmd = np.array([[float(Lm[i, j]) for j in range(n)] for i in range(n)])
    assert np.abs(np.tril(Lq) - Lmd).max() <= 2 * eps_of(np.float64) * np.abs(Lmd).max()


def test_quad_is_closer_to_truth_than_double(rng):
    n = 200
    A = spd(n, rng)
    Ld, _ = O.chol_recursive(A)
    Lq, _ = O.chol_recursive_q(A)
    assert O.chol_residual_q(A, np.tril(Lq)) <= O.chol_residual_q(A, np.tril(Ld))
    assert O.chol_residual_q(A, np.tril(Lq)) < 1e-16


@pytest.mark.parametrize("n,k", [(7, 3), (40, 1), (33, 17)])
def test_rank_updates_generic(rng, n, k):
    """rankUpdate! rank-1 and rank-k touch only the lower triangle."""
    C = rng.standard_normal((n, n))
    A = rng.standard_normal((n, k))
    out = O.syrk_l(C, A, -0.5)
    assert rel(np.tril(out), np.tril(C - 0.5 * A @ A.T)) <= 10 * n * eps_of(np.float64)
    assert np.array_equal(np.triu(out, 1), np.triu(C, 1))
    x = A[:, 0]
    out1 = O.syr_l(C, x, 2.0)
    assert rel(np.tril(out1), np.tril(C + 2.0 * np.outer(x, x))) <= 10 * n * eps_of(np.float64)


def test_trsm_matches_scipy(rng):
    L = np.linalg.cholesky(spd(40, rng))
    B = rng.standard_normal((25, 40))
    X = O.trsm_rlt(B, L)
    ref = sl.solve_triangular(L, B.T, lower=True).T
    assert rel(X, ref) <= 400 * eps_of(np.float64)


# ------------------------------------------------------------------- QR ----
def test_reflector_known_answer():
    x, tau = O.reflector(np.array([3.0, 4.0]))
    assert x[0] == -5.0 and x[1] == 0.5 and tau == 1.6


def test_reflector_degenerate_cases():
    x, tau = O.reflector(np.array([2.5, 0.0, 0.0]))
    assert tau == 2.0 and x[0] == -2.5          # Julia's τ = 2 (LAPACK dlarfg gives 0)
    x, tau = O.reflector(np.array([-1.5]))
    assert tau == 2.0 and x[0] == 1.5           # length-1: also τ = 2
    x, tau = O.reflector(np.zeros(4))
    assert tau == 0.0 and np.all(x == 0)


def test_reflector_matches_lapack_dlarfg(rng):
    """Non-degenerate: the same β = −sign(x₁)‖x‖, v, τ as LAPACK dlarfg."""
    x0 = rng.standard_normal(9)
    x, tau = O.reflector(x0)
    beta, v, tl = lapack.dlarfg(9, x0[0], x0[1:].copy())
    assert abs(x[0] - beta) <= 4 * eps_of(np.float64) * abs(beta)
    assert rel(x[1:], v) <= 8 * eps_of(np.float64)
    assert abs(tau - tl) <= 4 * eps_of(np.float64)


@pytest.mark.parametrize("m,n", [(10, 5), (10, 10), (5, 10), (100, 50), (100, 100), (50, 100)])
def test_qr_unblocked_matches_dgeqrf(rng, m, n):
    A = rng.standard_normal((m, n))
    F, tau = O.qr_unblocked(A)
    qr, tl, _, info = lapack.dgeqrf(A)
    k = min(m, n)
    tl = tl[:k].copy()
    # LAPACK's last τ of a wide/square problem is 0 where Julia's is 2 (1-element reflector)
    if m <= n:
        assert tau[m - 1] == 2.0 and tl[m - 1] == 0.0
        qr[m - 1, m - 1] *= -1
        qr[m - 1, m:] *= -1
        tl[m - 1] = 2.0
    assert rel(F, qr) <= 10 * max(m, n) * eps_of(np.float64)
    assert rel(tau, tl) <= 10 * max(m, n) * eps_of(np.float64)


@pytest.mark.parametrize("m,n", [(10, 5), (5, 10), (100, 50), (50, 100), (97, 97)])
@pytest.mark.parametrize("bz", [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33])
def test_qr_blocked_is_blocksize_invariant(m, n, bz):
    rng = np.random.default_rng(m * 100 + n + bz)
    A = rng.standard_normal((m, n))
    F0, t0 = O.qr_unblocked(A)
    F, t = O.qr_blocked(A, bz)
    assert rel(F, F0) <= 10 * max(m, n) * eps_of(np.float64)
    assert rel(t, t0) <= 10 * max(m, n) * eps_of(np.float64)


@pytest.mark.parametrize("m,n", [(12, 7), (7, 12), (30, 30)])
def test_tblock_equals_reflector_product(rng, m, n):
    A = rng.standard_normal((m, n))
    F, tau = O.qr_unblocked(A)
    k = min(m, n)
    T = O.qr_tblock(F, tau)
    V = np.tril(F[:, :k], -1) + np.eye(m, k)
    Q = np.eye(m)
    for j in range(k):
        Q = Q @ (np.eye(m) - tau[j] * np.outer(V[:, j], V[:, j]))
    assert np.abs(np.eye(m) - V @ T @ V.T - Q).max() <= 50 * m * eps_of(np.float64)
    assert np.array_equal(T, np.triu(T))


def test_tblock_matches_dlarft_recurrence(rng):
    """The UT transform and LAPACK's dlarft forward recurrence agree."""
    m, n = 40, 16
    F, tau = O.qr_unblocked(rng.standard_normal((m, n)))
    V = np.tril(F, -1) + np.eye(m, n)
    Tr = np.zeros((n, n))
    for j in range(n):
        Tr[:j, j] = -tau[j] * Tr[:j, :j] @ (V[:, :j].T @ V[:, j])
        Tr[j, j] = tau[j]
    assert rel(O.qr_tblock(F, tau), Tr) <= 100 * eps_of(np.float64)


@pytest.mark.parametrize("adj", [False, True])
def test_wy_apply_equals_explicit_q(rng, adj):
    m, n, nA = 30, 9, 11
    F, tau = O.qr_unblocked(rng.standard_normal((m, n)))
    T = O.qr_tblock(F, tau)
    V = np.tril(F, -1) + np.eye(m, n)
    Q = np.eye(m) - V @ T @ V.T
    B = rng.standard_normal((m, nA))
    out = O.wy_apply(F, T, B, adj)
    assert rel(out, (Q.T if adj else Q) @ B) <= 50 * m * eps_of(np.float64)


def test_reference_qr_invariants():
    """test/qr.jl:21-25 on the oracle: Q'A = R, Q'(QA) = A."""
    rng = np.random.default_rng(3)
    for (m, n) in [(10, 5), (5, 10), (100, 50)]:
        A = rng.standard_normal((m, n))
        F, tau = O.qr_blocked(A, 7)
        T = O.qr_tblock(F, tau)
        QtA = O.wy_apply(F, T, A, True)
        if m >= n:
            assert rel(QtA[:n], np.triu(F[:n])) <= 1e-13
        else:
            assert rel(QtA, np.triu(F)) <= 1e-13
        assert rel(O.wy_apply(F, T, O.wy_apply(F, T, A, False), True), A) <= 1e-13


def test_lmul_unit_lower(rng):
    L = rng.standard_normal((6, 6))
    B = rng.standard_normal((6, 4))
    out = O.lmul_unit_lower(L, B, 0.5)
    assert rel(out, 0.5 * (np.tril(L, -1) + np.eye(6)) @ B) <= 100 * eps_of(np.float64)


def test_qr_quad_close_to_double(rng):
    A = rng.standard_normal((60, 25))
    Fq, tq = O.qr_unblocked_q(A)
    Fd, td = O.qr_unblocked(A)
    assert rel(Fd, Fq) <= 600 * eps_of(np.float64)


# --------------------------------------------------------------- golden ----
def _golden(name):
    p = os.path.join(GOLDEN, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not generated")
    return np.load(p, allow_pickle=False)


def test_golden_cholesky():
    g = _golden("chol_golden.npz")
    for key in [k for k in g.files if k.startswith("A_")]:
        tag = key[2:]
        A, Lref = g[key], g["L_" + tag]
        n = A.shape[0]
        out, info = O.chol_recursive(A)
        assert info == 0
        assert rel(np.tril(out), Lref) <= 10 * n * eps_of(np.float64), tag


def test_golden_qr():
    g = _golden("qr_golden.npz")
    for key in [k for k in g.files if k.startswith("A_")]:
        tag = key[2:]
        A = g[key]
        F, tau = O.qr_blocked(A, 16)
        assert rel(F, g["F_" + tag]) <= 10 * max(A.shape) * eps_of(np.float64), tag
        assert rel(tau, g["tau_" + tag]) <= 10 * max(A.shape) * eps_of(np.float64), tag


@pytest.mark.parametrize("m,bz", [(40, 8), (25, 12), (9, 5)])
def test_qr_blocked_work_closed_form(rng, m, bz):
    """With n = 2·blocksize there is one lmul! (src/qr.jl:132) and its
    workspace ends as M = −V₁Tᵀ(VᵀA₂), whose top-k rows satisfy
    A₁ + M = R₁₂ (src/householder.jl:153-154): work = R₁₂ − A₁ with A₁ the
    original rows 0:k of columns bz:2bz and R₁₂ the final R there.  A
    transposed V₁, T in place of Tᵀ or a dropped axpy! fails this."""
    n = 2 * bz
    A = rng.standard_normal((m, n))
    F, tau, work = O.qr_blocked(A, bz, return_work=True)
    k = min(m, bz)
    expect = F[:k, bz:n] - A[:k, bz:n]
    assert np.abs(work[:k, :n - bz] - expect).max() <= 1e-13 * np.abs(A).max()
    # rows the step never writes keep their initial zeros
    assert not work[k:, :].any() and not work[:, n - bz:].any()
