"""The benchmark configurations at their full sizes, checked through
size-independent properties where the oracle is too slow, and per matrix
against the oracle where it is fast enough:

- batched uneven Cholesky + solve: 10,000 matrices, n ~ U[8, 256], FP64 and
  FP32, every L against the oracle's cholUnblocked! (src/cholesky.jl:3-15)
  and every x against a float64 solve, with few workgroups so that each
  takes many items of both kinds, and through a captured graph;
- cholRecursive! n = 16384: ‖A − LLᵀ‖/‖A‖ on sampled rows ≤ 10·n·eps, upper
  triangle untouched (src/cholesky.jl:37-55);
- qrBlocked! 8192×4096 at block sizes 64 and 128: QᵀA = R, orthonormal Q
  columns, and the same factors at both block sizes (src/qr.jl:113-146,
  test/qr.jl:21-25);
- the adjoints at their benchmark shapes: ⟨Ā, dA⟩ against a float64
  central difference along one random direction (Cholesky n = 8192 FP32,
  QR 4096×256 FP64), and 512 matrices of order 128 against the oracle.
"""
from concurrent.futures import ThreadPoolExecutor

from threadpoolctl import threadpool_limits

import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS64 = eps_of(np.float64)


def _gla():
    import gla

    return gla


# ------------------------------------------------------------ batched C4 --
def _packed_batch(gla, sizes, dtype, seed):
    """A PackedBatch of A_i = B_iᵀB_i + n_i·I and b_i ~ N(0,1), filled with
    one host-to-device copy; returns it with the host copies."""
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rng = np.random.default_rng(seed)
    pb = gla.PackedBatch(sizes, tdt)
    Ah = np.zeros(pb.A.numel(), dtype=dtype)
    bh = np.zeros(pb.b.numel(), dtype=dtype)
    As, bs = [], []
    for i, n in enumerate(sizes):
        B = rng.standard_normal((n, n))
        A = (B.T @ B + n * np.eye(n)).astype(dtype)
        b = rng.standard_normal(n).astype(dtype)
        Ah[pb.aoff[i]:pb.aoff[i] + n * n] = A.ravel(order="F")
        bh[pb.boff[i]:pb.boff[i] + n] = b
        As.append(A)
        bs.append(b)
    pb.A.copy_(torch.from_numpy(Ah))
    pb.b.copy_(torch.from_numpy(bh))
    return pb, As, bs, Ah, bh


def _check_batch(pb, As, bs, dtype):
    """Every L against the oracle's cholUnblocked! and every x against a
    float64 solve; the strict upper triangles keep their input."""
    torch.cuda.synchronize()
    Ag = pb.A.cpu().numpy()
    bg = pb.b.cpu().numpy()
    assert int(pb.info.abs().sum()) == 0

    def one(i):
        A = As[i]
        n = A.shape[0]
        M = Ag[pb.aoff[i]:pb.aoff[i] + n * n].reshape((n, n), order="F")
        Lo, info = O.chol_unblocked(A)
        if info != 0:
            return f"{i}: oracle info {info}"
        tol = 10 * n * eps_of(dtype)
        e = rel(np.tril(M), np.tril(Lo))
        if e > tol:
            return f"{i} (n={n}): L differs by {e:.2e} > {tol:.2e}"
        if not np.array_equal(np.triu(M, 1), np.triu(A, 1)):
            return f"{i} (n={n}): strict upper triangle changed"
        xr = np.linalg.solve(A.astype(np.float64), bs[i].astype(np.float64))
        ex = rel(bg[pb.boff[i]:pb.boff[i] + n], xr)
        if ex > tol * 10:        # κ(A) ≤ 10 for A = BᵀB + n·I
            return f"{i} (n={n}): x differs by {ex:.2e}"
        return None

    # one BLAS thread per worker: 16 workers each driving a 16-thread
    # OpenBLAS oversubscribe the host and can stall for minutes
    with threadpool_limits(1), ThreadPoolExecutor(16) as ex:
        errs = [e for e in ex.map(one, range(len(As))) if e]
    assert not errs, errs[:5]


@pytest.mark.parametrize("dtype,wave_max,grid_mult", [
    (np.float64, None, None),      # the autotuned bins
    (np.float64, 128, 1),          # both item kinds, ~10 items per workgroup
    (np.float32, None, None),
    (np.float32, 64, 1),
])
def test_batched_10000_uneven_vs_oracle(dtype, wave_max, grid_mult):
    gla = _gla()
    seed = 4000 + (0 if wave_max is None else wave_max) + (1 if dtype == np.float32 else 0)
    sizes = np.random.default_rng(seed).integers(8, 257, size=10000)
    pb, As, bs, _, _ = _packed_batch(gla, sizes, dtype, seed + 1)
    plan = gla.BatchPlan(sizes, pb.dtype, wave_max=wave_max, grid_mult=grid_mult)
    gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    _check_batch(pb, As, bs, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_10000_graph_replays(dtype):
    """The benchmark's form: one captured graph, replayed twice with the
    inputs restored eagerly in between, every replay checked."""
    gla = _gla()
    sizes = np.random.default_rng(77).integers(8, 257, size=10000)
    pb, As, bs, Ah, bh = _packed_batch(gla, sizes, dtype, 78)
    A0, b0 = pb.A.clone(), pb.b.clone()
    plan = gla.BatchPlan(sizes, pb.dtype)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    for _ in range(2):
        pb.A.copy_(A0)
        pb.b.copy_(b0)
        pb.info.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _check_batch(pb, As, bs, dtype)


# ------------------------------------------------------- Cholesky n=16384 --
def _spd_device(n, seed, dtype=torch.float64):
    g = torch.Generator(device="cuda").manual_seed(seed)
    B = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A = B.T @ B
    A.diagonal().add_(n)
    return A.to(dtype)


@pytest.mark.parametrize("chol_block", [-1, 768, 0])
def test_cholesky_16384_residual_on_sampled_rows(chol_block):
    gla = _gla()
    n = 16384
    A = _spd_device(n, 1234)                 # symmetric: row- and column-major alike
    W = A.clone()                            # column-major view of the same values
    with gla.tuning(chol_block=chol_block):
        gla.cholRecursive_(W.t(), gla.Val.L)
    # W.t() is column-major: element (i, j) of the factor is W[j, i]
    L = torch.tril(W.t())
    rows = torch.from_numpy(np.sort(np.random.default_rng(5).choice(n, 96, replace=False)))
    rows = rows.to("cuda")
    R = A[rows] - L[rows] @ L.T
    err = float(torch.linalg.norm(R) / torch.linalg.norm(A[rows]))
    assert err <= 10 * n * EPS64, err
    # the strict upper triangle (of the column-major matrix) keeps its input
    assert torch.equal(torch.triu(W.t(), 1), torch.triu(A, 1))
    assert bool(torch.all(torch.diagonal(L) > 0))


# ----------------------------------------------------- QR 8192×4096 -------
def test_qr_8192x4096_blocksize_64_and_128():
    gla = _gla()
    m, n = 8192, 4096
    g = torch.Generator(device="cuda").manual_seed(99)
    A = torch.randn(n, m, dtype=torch.float64, device="cuda", generator=g).t()  # column-major
    F128 = gla.qrBlocked_(gla.colmajor(A, device="cuda"), 128)
    F64 = gla.qrBlocked_(gla.colmajor(A, device="cuda"), 64)
    tol = 10 * m * EPS64
    d = float(torch.linalg.norm(F128.factors - F64.factors) / torch.linalg.norm(F64.factors))
    assert d <= tol, d
    dt = float(torch.linalg.norm(F128.tau - F64.tau) / torch.linalg.norm(F64.tau))
    assert dt <= tol, dt
    # QᵀA = R through the compact-WY apply with the whole factorisation's T
    H = F128[gla.QBlocked]
    QtA = gla.lmul_(gla.adjoint(H), gla.colmajor(A, device="cuda"))
    R = torch.triu(F128.factors[:n])
    e1 = float(torch.linalg.norm(QtA[:n] - R) / torch.linalg.norm(R))
    e2 = float(torch.linalg.norm(QtA[n:]) / torch.linalg.norm(A))
    assert e1 <= tol and e2 <= tol, (e1, e2)
    # orthonormal columns of Q on a sample
    cols = np.sort(np.random.default_rng(3).choice(m, 64, replace=False))
    E = torch.zeros(m, 64, dtype=torch.float64, device="cuda")
    E[torch.from_numpy(cols).cuda(), torch.arange(64, device="cuda")] = 1.0
    X = gla.lmul_(H, gla.colmajor(E, device="cuda"))
    eo = float(torch.linalg.norm(X.T @ X - torch.eye(64, dtype=torch.float64, device="cuda")))
    assert eo <= tol, eo


# ------------------------------------------------------------- adjoints C5 --
def _dot(a, b):
    return float((a.to(torch.float64) * b.to(torch.float64)).sum())


def test_cholesky_adjoint_8192_fp32_directional():
    """⟨Ā, dA⟩ (FP32 adjoint of the FP32 factor, symmetric convention) equals
    d/dt ⟨W, L(A + t·dA)⟩ from a float64 central difference (t = 1e-6,
    float64 factorisations); bar 1e-3 relative, the FP32 adjoint's accuracy
    at n = 8192 (n·eps32 = 4.9e-4)."""
    gla = _gla()
    n = 8192
    A = _spd_device(n, 21)
    g = torch.Generator(device="cuda").manual_seed(22)
    W = torch.tril(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    D = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    dA = (D + D.T) * (float(torch.linalg.norm(A)) / float(torch.linalg.norm(D + D.T)))

    def f(t):
        M = (A + t * dA).clone()
        gla.cholRecursive_(M.t(), gla.Val.L)
        return _dot(W, torch.tril(M.t()))

    h = 1e-6
    fd = (f(h) - f(-h)) / (2 * h)
    L32 = gla.colmajor(A.t(), torch.float32)
    gla.cholRecursive_(L32, gla.Val.L)
    Abar = gla.chol_adjoint(L32, gla.colmajor(W, torch.float32), "S")
    got = _dot(Abar, dA)
    assert abs(got - fd) <= 1e-3 * abs(fd), (got, fd)


def test_qr_adjoint_4096x256_directional():
    """⟨Ā, dA⟩ for R̄ = triu(W) against the central difference of ⟨W, R(A)⟩
    (t = 1e-6, float64); bar 1e-6 relative."""
    gla = _gla()
    m, n = 4096, 256
    g = torch.Generator(device="cuda").manual_seed(31)
    A = torch.randn(m, n, dtype=torch.float64, device="cuda", generator=g)
    W = torch.triu(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    dA = torch.randn(m, n, dtype=torch.float64, device="cuda", generator=g)

    def f(t):
        q = gla.qrBlocked_(gla.colmajor(A + t * dA), 128)
        return _dot(W, torch.triu(q.factors[:n]))

    h = 1e-6
    fd = (f(h) - f(-h)) / (2 * h)
    q = gla.qrBlocked_(gla.colmajor(A), 128)
    Abar = gla.qr_adjoint(q.factors, q.tau, gla.colmajor(W))
    got = _dot(Abar, dA)
    assert abs(got - fd) <= 1e-6 * abs(fd), (got, fd)


def test_cholesky_adjoint_batched_512x128_vs_oracle():
    """512 matrices of order 128 (factors recomputed on the device) against
    the oracle's adjoint of each matrix's LAPACK factor, 50·n·eps."""
    gla = _gla()
    rng = np.random.default_rng(512)
    n, batch = 128, 512
    As = np.empty((batch, n, n))
    for b in range(batch):
        B = rng.standard_normal((n, n))
        As[b] = B.T @ B + n * np.eye(n)
    Ws = np.tril(rng.standard_normal((batch, n, n)))
    At = torch.from_numpy(As).cuda().transpose(1, 2)   # symmetric, so values agree
    Wt = torch.from_numpy(np.ascontiguousarray(Ws.transpose(0, 2, 1))).cuda().transpose(1, 2)
    got = gla.chol_adjoint_batched(At, Wt, "S").cpu().numpy()

    def one(b):
        ref = O.chol_adjoint(np.linalg.cholesky(As[b]), Ws[b], "S")
        e = rel(got[b], ref)
        return None if e <= 50 * n * EPS64 else f"{b}: {e:.2e}"

    # one BLAS thread per worker: 16 workers each driving a 16-thread
    # OpenBLAS oversubscribe the host and can stall for minutes
    with threadpool_limits(1), ThreadPoolExecutor(16) as ex:
        errs = [e for e in ex.map(one, range(batch)) if e]
    assert not errs, errs[:5]
