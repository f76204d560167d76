"""Batched uneven Cholesky + solve (one persistent launch) against the
oracle's cholUnblocked! (src/cholesky.jl:3-15) per matrix and a float64
solve.  Bar: normwise relative difference ≤ 10·n·eps for L and for x."""
import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def _fill(gla, sizes, dtype, rng, bad=None):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    pb = gla.PackedBatch(sizes, tdt)
    As, bs = [], []
    for i, n in enumerate(sizes):
        B = rng.standard_normal((n, n))
        A = B.T @ B + n * np.eye(n)
        if bad is not None and i in bad:
            k = bad[i]
            Lr = np.linalg.cholesky(A)
            A[k, k] -= Lr[k, k] ** 2 + 1.0
        A = A.astype(dtype)
        b = rng.standard_normal(n).astype(dtype)
        pb.matrix(i).copy_(torch.from_numpy(A))
        pb.vector(i).copy_(torch.from_numpy(b))
        As.append(A)
        bs.append(b)
    return pb, As, bs


def _check(pb, As, bs, dtype):
    torch.cuda.synchronize()
    for i, (A, b) in enumerate(zip(As, bs)):
        n = A.shape[0]
        if n == 0:
            continue
        L = np.tril(pb.matrix(i).cpu().numpy())
        x = pb.vector(i).cpu().numpy()
        Lo, info = O.chol_unblocked(A)
        assert info == 0
        tol = 10 * n * eps_of(dtype)
        assert rel(L, np.tril(Lo)) <= tol, (i, n)
        xr = np.linalg.solve(A.astype(np.float64), b.astype(np.float64))
        assert rel(x, xr) <= tol * max(1.0, np.linalg.cond(A.astype(np.float64))), (i, n)
        # strict upper triangle keeps the input
        Ain = As[i]
        up = np.triu(pb.matrix(i).cpu().numpy(), 1)
        assert np.array_equal(up, np.triu(Ain, 1)), (i, n)


EDGE = [1, 2, 7, 8, 16, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 160, 200, 255, 256]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("wave_max", [0, 32, 128, 256])
def test_uneven_sizes_vs_oracle(dtype, wave_max):
    gla = _gla()
    rng = np.random.default_rng(wave_max + (1 if dtype == np.float32 else 0))
    sizes = EDGE + list(rng.integers(8, 257, size=40))
    rng.shuffle(sizes)
    pb, As, bs = _fill(gla, sizes, dtype, rng)
    plan = gla.BatchPlan(sizes, pb.dtype, wave_max=wave_max, grid_mult=2)
    gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    _check(pb, As, bs, dtype)
    assert int(pb.info.abs().sum()) == 0


def test_negative_pivot_reported_per_matrix():
    gla = _gla()
    rng = np.random.default_rng(11)
    sizes = [20, 150, 40, 256, 9]
    bad = {1: 77, 3: 200, 4: 0}
    pb, As, bs = _fill(gla, sizes, np.float64, rng, bad=bad)
    plan = gla.BatchPlan(sizes, pb.dtype)
    gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    info = pb.info.cpu().numpy()
    for i in range(len(sizes)):
        assert info[i] == (bad[i] + 1 if i in bad else 0), (i, info)
        _, oi = O.chol_unblocked(As[i])
        assert info[i] == oi


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_graph_capture_replays(dtype):
    """hipGraph capture of the batched call: replayed three times, inputs
    restored eagerly between replays, every replay checked."""
    gla = _gla()
    rng = np.random.default_rng(5)
    sizes = list(rng.integers(8, 257, size=64))
    pb, As, bs = _fill(gla, sizes, dtype, rng)
    A0, b0 = pb.A.clone(), pb.b.clone()
    plan = gla.BatchPlan(sizes, pb.dtype)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):                       # warm the handle on this stream
        gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    pb.A.copy_(A0)
    pb.b.copy_(b0)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        gla.chol_solve_batched_(plan, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    for _ in range(3):
        pb.A.copy_(A0)
        pb.b.copy_(b0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _check(pb, As, bs, dtype)


def test_empty_and_zero_size():
    gla = _gla()
    plan = gla.BatchPlan([], torch.float64)
    pb = gla.PackedBatch([0, 3], torch.float64)
    pb.matrix(1).copy_(torch.eye(3, dtype=torch.float64) * 4)
    pb.vector(1).copy_(torch.ones(3, dtype=torch.float64))
    plan2 = gla.BatchPlan([0, 3], torch.float64)
    gla.chol_solve_batched_(plan2, pb.A_ptrs, pb.lda, pb.b_ptrs, pb.info)
    torch.cuda.synchronize()
    assert np.allclose(pb.vector(1).cpu().numpy(), 0.25)
    assert np.allclose(np.diag(pb.matrix(1).cpu().numpy()), 2.0)
    del plan


def test_size_out_of_range_rejected():
    gla = _gla()
    with pytest.raises(RuntimeError, match="outside"):
        gla.BatchPlan([10, 300], torch.float64)
