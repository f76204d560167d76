"""Householder QR on the device against the oracle (restatement of
src/qr.jl and src/householder.jl) and the reference's invariants
(test/qr.jl:6-36).  Bar: normwise relative difference ≤ 10·n·eps for the
packed factors (R and V), τ, and T."""
import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def _tdt(dtype):
    return torch.float64 if dtype == np.float64 else torch.float32


DIMS = [(10, 5), (10, 10), (5, 10), (100, 50), (100, 100), (50, 100)]
BZS = [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 31, 32, 33]


@pytest.mark.parametrize("m,n", DIMS)
@pytest.mark.parametrize("bz", BZS)
def test_reference_grid(m, n, bz):
    """test/qr.jl:7-26 — Q'A = R (m ≥ n) or triu(factors) (m < n), Q'(QA) = A —
    plus the factors themselves against the oracle's qrBlocked!."""
    gla = _gla()
    rng = np.random.default_rng(m * 1000 + n * 10 + bz)
    A = rng.standard_normal((m, n))
    F = gla.qrBlocked_(gla.colmajor(A), bz)
    fac, tau = gla.to_numpy(F.factors), F.tau.cpu().numpy()
    ofac, otau = O.qr_blocked(A, bz)
    k = max(m, n)
    tol = 10 * k * eps_of(np.float64)
    assert rel(fac, ofac) <= tol
    assert rel(tau, otau) <= tol
    Q = F[gla.QBlocked]
    QtA = gla.to_numpy(gla.adjoint(Q) @ gla.colmajor(A))
    if m >= n:
        Rm = gla.to_numpy(F[gla.R].data)
        assert rel(QtA[:min(m, n), :], np.triu(Rm)) <= tol
    else:
        with pytest.raises(gla.ArgumentError):
            F[gla.R]
        assert rel(QtA, np.triu(fac)) <= tol
    QQA = gla.to_numpy(gla.adjoint(Q) @ (Q @ gla.colmajor(A)))
    assert rel(QQA, A) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (7, 1), (64, 64), (129, 128), (300, 128),
                                 (1000, 257), (4100, 300), (20000, 256)])
def test_unblocked_and_blocked_vs_oracle(rng, dtype, m, n):
    gla = _gla()
    A = rng.standard_normal((m, n)).astype(dtype)
    tol = 10 * max(m, n) * eps_of(dtype)
    for nb in (None, 32, 128):
        Ag = gla.colmajor(A, _tdt(dtype))
        F = gla.qrUnblocked_(Ag) if nb is None else gla.qrBlocked_(Ag, nb)
        fac, tau = gla.to_numpy(F.factors), F.tau.cpu().numpy()
        ofac, otau = O.qr_blocked(A.astype(np.float64), nb or min(n, 128))
        assert rel(fac, ofac) <= tol, nb
        assert rel(tau, otau) <= tol, nb


@pytest.mark.parametrize("la", [0, 1])
@pytest.mark.parametrize("inner", [0, 32, 64])
@pytest.mark.parametrize("m,n,nb", [(300, 128, 128), (1000, 257, 128), (2000, 700, 100),
                                    (130, 400, 128), (5000, 192, 64), (97, 97, 128),
                                    (1500, 1500, 128)])
def test_subpanel_widths_and_lookahead_agree(rng, la, inner, m, n, nb):
    """Factoring each block column as sub-panels (qr_inner 32/64) or in one
    panel launch (0), with or without the look-ahead driver, computes the
    same Householder sequence."""
    gla = _gla()
    A = rng.standard_normal((m, n))
    tol = 10 * max(m, n) * eps_of(np.float64)
    ofac, otau = O.qr_blocked(A, nb)
    try:
        gla.set_tuning("qr_inner", inner)
        gla.set_tuning("qr_lookahead", la)
        F = gla.qrBlocked_(gla.colmajor(A), nb)
    finally:
        gla.set_tuning("qr_inner", -1)
        gla.set_tuning("qr_lookahead", 1)
    assert rel(gla.to_numpy(F.factors), ofac) <= tol
    assert rel(F.tau.cpu().numpy(), otau) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("la", [0, 1])
@pytest.mark.parametrize("m,n,nb", [(300, 128, 128), (1000, 257, 128), (2000, 700, 100),
                                    (129, 128, 128), (5000, 192, 64), (97, 97, 16),
                                    (1500, 1500, 128), (400, 3, 1), (700, 200, 33)])
def test_reconstructed_panel_vs_oracle(rng, dtype, la, m, n, nb):
    """Every block column through the reconstructed-Householder panel
    (qr_hr = 1: from one row up): CholeskyQR twice, then the signed LU of
    the orthonormal factor's top block.  Same V, τ, R as qrBlocked!."""
    gla = _gla()
    A = rng.standard_normal((m, n)).astype(dtype)
    tol = 10 * max(m, n) * eps_of(dtype)
    ofac, otau = O.qr_blocked(A.astype(np.float64), nb)
    try:
        gla.set_tuning("qr_hr", 1)
        gla.set_tuning("qr_lookahead", la)
        F = gla.qrBlocked_(gla.colmajor(A, _tdt(dtype)), nb)
        QtA = gla.to_numpy(gla.adjoint(F[gla.QBlocked]) @ gla.colmajor(A, _tdt(dtype)))
    finally:
        gla.set_tuning("qr_hr", -1)
        gla.set_tuning("qr_lookahead", 1)
    assert rel(gla.to_numpy(F.factors), ofac) <= tol
    assert rel(F.tau.cpu().numpy(), otau) <= tol
    k = min(m, n)
    assert rel(QtA[:k, :], np.triu(ofac[:k, :])) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reconstructed_panel_full_second_cholesky(rng, dtype):
    """A first block column with κ = 3e4 (FP64; 1e2 in FP32): the first
    CholeskyQR pass leaves ‖Q₁ᵀQ₁ − I‖ above 0.1·√eps (checked here on the
    host), so the second pass takes the full 128×128 Cholesky rather than
    the first-order factor of I + E.  Same V, τ, R as qrBlocked!."""
    gla = _gla()
    m, n, nb = 3000, 160, 64
    kappa = 3e4 if dtype == np.float64 else 1e2
    A = rng.standard_normal((m, n))
    U, _ = np.linalg.qr(rng.standard_normal((m, nb)))
    W, _ = np.linalg.qr(rng.standard_normal((nb, nb)))
    A[:, :nb] = (U * np.logspace(0, -np.log10(kappa), nb)) @ W.T
    A = A.astype(dtype)
    P = A[:, :nb]
    L = np.linalg.cholesky((P.T @ P).astype(np.float64)).astype(dtype)
    Q1 = (P @ np.linalg.inv(L.T.astype(np.float64)).astype(dtype)).astype(np.float64)
    E = np.linalg.norm(Q1.T @ Q1 - np.eye(nb))
    assert E > 0.2 * np.sqrt(eps_of(dtype)), E
    tol = 10 * max(m, n) * eps_of(dtype)
    ofac, otau = O.qr_blocked(A.astype(np.float64), nb)
    try:
        gla.set_tuning("qr_hr", 1)
        F = gla.qrBlocked_(gla.colmajor(A, _tdt(dtype)), nb)
    finally:
        gla.set_tuning("qr_hr", -1)
    assert rel(gla.to_numpy(F.factors), ofac) <= tol
    assert rel(F.tau.cpu().numpy(), otau) <= tol


@pytest.mark.parametrize("case", ["zero_column", "rank_one", "graded"])
def test_reconstructed_panel_falls_back(rng, case):
    """Block columns the Cholesky-based path cannot factor to working
    accuracy — a zero column (τ = 0 in reflector!), rank deficiency, columns
    graded over 14 orders of magnitude — fail its device checks and are
    factored column by column instead: the result is still qrBlocked!'s."""
    gla = _gla()
    m, n, nb = 3000, 160, 64
    A = rng.standard_normal((m, n))
    if case == "zero_column":
        A[:, 5] = 0.0
    elif case == "rank_one":
        A[:, 64:128] = np.outer(A[:, 64], rng.standard_normal(64))
    else:
        A[:, :nb] *= np.logspace(0, -14, nb)[None, :]
    tol = 10 * max(m, n) * eps_of(np.float64)
    ofac, otau = O.qr_blocked(A, nb)
    try:
        gla.set_tuning("qr_hr", 1)
        F = gla.qrBlocked_(gla.colmajor(A), nb)
    finally:
        gla.set_tuning("qr_hr", -1)
    fac, tau = gla.to_numpy(F.factors), F.tau.cpu().numpy()
    assert np.all(np.isfinite(fac)) and np.all(np.isfinite(tau))
    if case == "zero_column":
        assert tau[5] == 0.0
    if case == "rank_one":
        # trailing columns of a rank-deficient block are rounding noise: compare
        # the well-determined part and the factorisation itself
        assert rel(fac[:, :65], ofac[:, :65]) <= tol
        Q = F[gla.QBlocked]
        QtA = gla.to_numpy(gla.adjoint(Q) @ gla.colmajor(A))
        assert np.linalg.norm(np.tril(QtA, -1)) <= tol * np.linalg.norm(A)
        return
    assert rel(fac, ofac) <= tol
    assert rel(tau, otau) <= tol


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_reconstructed_panel_non_finite_input(rng, bad):
    """A NaN or Inf in a tall block column makes the Gram matrices
    non-finite: the device check flags the reconstructed path and the
    column-by-column kernel factors the block column, propagating the value
    as reflector! does (no error, no hang); the untouched block columns'
    reflectors before it match the oracle."""
    gla = _gla()
    m, n, nb = 3000, 256, 128
    A = rng.standard_normal((m, n))
    A[1234, 200] = bad                      # second block column
    ofac, otau = O.qr_blocked(A, nb)
    F = gla.qrBlocked_(gla.colmajor(A), nb)
    fac, tau = gla.to_numpy(F.factors), F.tau.cpu().numpy()
    assert not np.all(np.isfinite(fac[:, 128:]))
    tol = 10 * m * eps_of(np.float64)
    # the first block column's reflectors and τ are computed before the bad
    # value can reach them
    assert rel(fac[128:, :128], ofac[128:, :128]) <= tol
    assert rel(tau[:128], otau[:128]) <= tol


def test_known_reflector():
    """[3;4]: β = −5, v = [1, 0.5], τ = 1.6 (Julia's reflector! convention)."""
    gla = _gla()
    F = gla.qrUnblocked_(gla.colmajor(np.array([[3.0], [4.0]])))
    fac = gla.to_numpy(F.factors)
    assert np.allclose(fac[:, 0], [-5.0, 0.5], rtol=0, atol=1e-15)
    assert abs(float(F.tau[0]) - 1.6) < 1e-15


def test_reduced_column_gives_tau_two():
    """A column [a, 0, …] with a ≠ 0 gives τ = 2 and β = −a (unlike LAPACK's τ = 0);
    a zero column gives τ = 0 and stays zero."""
    gla = _gla()
    A = np.zeros((6, 3))
    A[0, 0] = 2.5
    A[:, 2] = np.arange(1, 7)
    F = gla.qrUnblocked_(gla.colmajor(A))
    fac, tau = gla.to_numpy(F.factors), F.tau.cpu().numpy()
    ofac, otau = O.qr_unblocked(A)
    assert tau[0] == 2.0 and fac[0, 0] == -2.5
    assert tau[1] == 0.0
    assert rel(fac, ofac) <= 1e-14 and rel(tau, otau) <= 1e-14


@pytest.mark.parametrize("m,n", [(40, 17), (17, 40), (300, 200), (600, 600)])
def test_tblock_matches_explicit_reflector_product(rng, m, n):
    """F[Tuple{:QBlocked}]: I − V T Vᵀ equals H₁H₂⋯H_k formed explicitly."""
    gla = _gla()
    A = rng.standard_normal((m, n))
    F = gla.qrBlocked_(gla.colmajor(A), 16)
    H = F[gla.QBlocked]
    T = gla.to_numpy(H.T.data)
    fac, tau = gla.to_numpy(F.factors), F.tau.cpu().numpy()
    k = min(m, n)
    Tref = O.qr_tblock(fac, tau)
    assert rel(np.triu(T), Tref) <= 10 * max(m, n) * eps_of(np.float64)
    if m <= 300:
        V = np.tril(fac[:, :k], -1) + np.eye(m, k)
        Q = np.eye(m)
        for j in range(k):
            Q = Q @ (np.eye(m) - tau[j] * np.outer(V[:, j], V[:, j]))
        assert np.abs(np.eye(m) - V @ np.triu(T) @ V.T - Q).max() <= 100 * m * eps_of(np.float64)


@pytest.mark.parametrize("adj", [False, True])
@pytest.mark.parametrize("mA,nV,nA", [(50, 8, 13), (300, 64, 100), (2000, 128, 700), (5, 8, 3)])
def test_wy_apply_vs_oracle(rng, adj, mA, nV, nA):
    """lmul!(H, A, M) / lmul!(H', A, M) — src/householder.jl:82-157."""
    gla = _gla()
    V = rng.standard_normal((mA, nV))
    b = min(mA, nV)
    T = np.triu(rng.standard_normal((b, b))) / np.sqrt(b)
    A = rng.standard_normal((mA, nA))
    H = gla.HouseholderBlock(gla.colmajor(V), gla.UpperTriangular(gla.colmajor(T)))
    Ag = gla.colmajor(A)
    M = gla.colmajor(np.zeros((b, nA)))
    gla.lmul_(gla.adjoint(H) if adj else H, Ag, M)
    ref = O.wy_apply(V, T, A, adj)
    assert rel(gla.to_numpy(Ag), ref) <= 10 * mA * eps_of(np.float64)


def test_wy_dimension_mismatch():
    gla = _gla()
    H = gla.HouseholderBlock(gla.colmajor(np.ones((5, 2))), gla.colmajor(np.eye(2)))
    with pytest.raises(gla.DimensionMismatch):
        gla.lmul_(H, gla.colmajor(np.ones((4, 3))))


@pytest.mark.parametrize("m,n", [(5, 5), (64, 30), (200, 77)])
def test_lmul_unit_lower(rng, m, n):
    """lmul!(UnitLowerTriangular(L), B, α) — src/juliaBLAS.jl:154-166."""
    gla = _gla()
    L = rng.standard_normal((m, m))
    B = rng.standard_normal((m, n))
    Bg = gla.colmajor(B)
    gla.lmul_(gla.UnitLowerTriangular(gla.colmajor(L)), Bg, -0.5)
    ref = O.lmul_unit_lower(L, B, -0.5)
    assert rel(gla.to_numpy(Bg), ref) <= 10 * m * eps_of(np.float64)


@pytest.mark.parametrize("m,n,bz", [(100, 50, 7), (100, 100, 16), (50, 100, 9), (10, 10, 3),
                                    (300, 200, 64), (700, 450, 128), (2000, 600, 128),
                                    (400, 300, 200)])
@pytest.mark.parametrize("la", [0, 1])
def test_work_left_as_the_reference_leaves_it(m, n, bz, la):
    """qrBlocked!'s `work` after the call (src/qr.jl:117,132): each step's
    lmul! workspace M = −V₁Tᵀ(VᵀA₂), later steps overwriting the top-left;
    compared with the oracle's work (blocksize > 128: as qrBlocked!(A, 128)
    leaves it, the documented block width)."""
    gla = _gla()
    rng = np.random.default_rng(m + 3 * n + bz)
    A = rng.standard_normal((m, n))
    obz = min(bz, 128)
    ofac, otau, owork = O.qr_blocked(A, obz, return_work=True)
    work = gla.colmajor(np.zeros((obz, n)))
    try:
        gla.set_tuning("qr_lookahead", la)
        F = gla.qrBlocked_(gla.colmajor(A), bz, work=work)
    finally:
        gla.set_tuning("qr_lookahead", 1)
    tol = 10 * max(m, n) * eps_of(np.float64)
    assert rel(gla.to_numpy(F.factors), ofac) <= tol
    assert rel(gla.to_numpy(work), owork) <= tol, rel(gla.to_numpy(work), owork)


def test_panel_wait_timeout_is_reported(rng):
    """A grid-wide wait of the column panel that gives up makes the call
    return GLA_ERR_TIMEOUT (DeviceTimeout): forced from the host by waits
    that target an arrival count that never comes, each giving up after a
    few polls, so every kernel still finishes.  The next call with the
    normal settings succeeds and matches the oracle."""
    gla = _gla()
    A = rng.standard_normal((4096, 128))
    try:
        gla.set_tuning("qr_hr", 0)           # the column panel: 32 workgroups
        gla.set_tuning("spin_limit", 64)
        gla.set_tuning("qr_wait_test", 1)
        with pytest.raises(gla.DeviceTimeout, match="gave up"):
            gla.qrBlocked_(gla.colmajor(A), 128)
    finally:
        gla.set_tuning("qr_wait_test", 0)
        gla.set_tuning("spin_limit", 1 << 25)
    try:
        F = gla.qrBlocked_(gla.colmajor(A), 128)
    finally:
        gla.set_tuning("qr_hr", -1)
    ofac, otau = O.qr_blocked(A, 128)
    tol = 10 * 4096 * eps_of(np.float64)
    assert rel(gla.to_numpy(F.factors), ofac) <= tol
    assert rel(F.tau.cpu().numpy(), otau) <= tol


def test_block_column_taller_than_the_64_wide_panel(rng):
    """70000 rows: more than the 64-wide panel's 256·256 rows, so the block
    column is factored as 32-wide sub-panels; against the oracle."""
    gla = _gla()
    m, n = 70000, 128
    A = rng.standard_normal((m, n))
    F = gla.qrBlocked_(gla.colmajor(A), 128)
    ofac, otau = O.qr_blocked(A, 128)
    tol = 10 * m * eps_of(np.float64)
    assert rel(gla.to_numpy(F.factors), ofac) <= tol
    assert rel(F.tau.cpu().numpy(), otau) <= tol


def test_block_column_beyond_capacity_is_refused_cleanly():
    """Above 512 rows per CU even 32-wide sub-panels cannot be resident at
    once: the call returns GLA_ERR_UNSUPPORTED with a message, nothing runs."""
    gla = _gla()
    A = gla.colmajor(np.ones((140000, 32)))
    with pytest.raises(RuntimeError, match="exceeds the register-resident panel"):
        gla.qrBlocked_(A, 32)


def test_masked_side_stream_is_bitwise_neutral(rng):
    """The look-ahead driver with its trailing updates on the CU-masked side
    stream and the panel chain on the handle's chain stream (the default)
    computes the same bits as with side_cu_reserve 0, also when called on a
    stream other than the default one."""
    gla = _gla()
    A = rng.standard_normal((3000, 1000))
    outs = []
    for reserve in (gla.get_tuning("side_cu_reserve"), 0):
        with gla.tuning(side_cu_reserve=reserve):
            X = gla.colmajor(A)
            st = torch.cuda.Stream()
            st.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(st):
                F = gla.qrBlocked_(X, 128)
            torch.cuda.current_stream().wait_stream(st)
            torch.cuda.synchronize()
            outs.append((gla.to_numpy(F.factors), F.tau.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    ofac, otau = O.qr_blocked(A, 128)
    assert rel(outs[0][0], ofac) <= 10 * 3000 * eps_of(np.float64)
