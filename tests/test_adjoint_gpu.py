"""Cholesky and QR adjoints on the device against the oracle's adjoints
(oracle.chol_adjoint / oracle.qr_adjoint, which the CPU suite pins against
PyTorch CPU autograd and central finite differences).  Bar: normwise
relative difference ≤ 50·n·eps of the element type (the adjoint applies
L⁻¹ twice; the inputs are well conditioned, κ(A) ≲ 5)."""
import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def _spd(n, rng):
    B = rng.standard_normal((n, n))
    return B.T @ B + n * np.eye(n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 2, 17, 64, 128, 129, 300, 520])
@pytest.mark.parametrize("conv", ["S", "L"])
def test_chol_adjoint_vs_oracle(dtype, n, conv):
    gla = _gla()
    rng = np.random.default_rng(n + (7 if conv == "L" else 0))
    A = _spd(n, rng)
    L = np.linalg.cholesky(A).astype(dtype)
    W = rng.standard_normal((n, n)).astype(dtype)
    # the strict upper parts of L and L̄ are ignored: fill them with junk
    Lj = np.tril(L) + np.triu(np.full((n, n), 7.0, dtype), 1)
    Wj = np.tril(W) + np.triu(np.full((n, n), -3.0, dtype), 1)
    got = gla.to_numpy(gla.chol_adjoint(gla.colmajor(Lj), gla.colmajor(Wj), conv))
    ref = O.chol_adjoint(L.astype(np.float64), W.astype(np.float64), conv)
    assert rel(got, ref) <= 50 * n * eps_of(dtype), rel(got, ref)
    if conv == "S":
        assert np.array_equal(got, got.T)
    else:
        assert not np.triu(got, 1).any()


def _batch_tensor(mats, tdt):
    """(batch, n, n) device tensor holding column-major matrices."""
    b, n = len(mats), mats[0].shape[0]
    T = torch.empty((b, n, n), dtype=tdt, device="cuda").transpose(1, 2)
    T.copy_(torch.from_numpy(np.stack(mats)).to(tdt))
    return T


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 16, 37, 128])
@pytest.mark.parametrize("conv", ["S", "L"])
def test_chol_adjoint_batched_vs_oracle(dtype, n, conv):
    """Batched form (factor recomputed from A) against the oracle applied to
    each matrix's LAPACK factor; strict upper parts of A and L̄ hold junk."""
    gla = _gla()
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rng = np.random.default_rng(100 + n + (7 if conv == "L" else 0))
    batch = 9
    As = [_spd(n, rng) for _ in range(batch)]
    Ws = [rng.standard_normal((n, n)) for _ in range(batch)]
    Aj = [np.tril(a) + np.triu(np.full((n, n), 1e3), 1) for a in As]
    Wj = [np.tril(w) + np.triu(np.full((n, n), -3.0), 1) for w in Ws]
    got = gla.chol_adjoint_batched(_batch_tensor(Aj, tdt), _batch_tensor(Wj, tdt), conv)
    got = got.cpu().numpy()
    for b in range(batch):
        L = np.linalg.cholesky(As[b].astype(dtype).astype(np.float64))
        ref = O.chol_adjoint(L, np.tril(Ws[b].astype(dtype).astype(np.float64)), conv)
        assert rel(got[b], ref) <= 50 * n * eps_of(dtype), (b, rel(got[b], ref))


def test_chol_adjoint_batched_reports_failing_matrix():
    gla = _gla()
    rng = np.random.default_rng(5)
    n, batch = 40, 6
    As = [_spd(n, rng) for _ in range(batch)]
    As[3][10, 10] = -1.0
    W = [rng.standard_normal((n, n)) for _ in range(batch)]
    with pytest.raises(gla.DomainError, match="matrix 3"):
        gla.chol_adjoint_batched(_batch_tensor(As, torch.float64), _batch_tensor(W, torch.float64))


def test_chol_adjoint_batched_rejects_large_order():
    gla = _gla()
    A = torch.eye(129, dtype=torch.float64, device="cuda").expand(2, 129, 129).contiguous()
    A = A.transpose(1, 2)
    with pytest.raises(RuntimeError):
        gla.chol_adjoint_batched(A, A.clone())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n", [(1, 1), (10, 6), (128, 128), (300, 129), (513, 200), (700, 256)])
def test_qr_adjoint_vs_oracle(dtype, m, n):
    gla = _gla()
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)).astype(dtype)
    q = gla.qrBlocked_(gla.colmajor(A), 64)
    F, tau = q.factors, q.tau
    W = rng.standard_normal((n, n)).astype(dtype)
    Wj = np.triu(W) + np.tril(np.full((n, n), 5.0, dtype), -1)
    got = gla.to_numpy(gla.qr_adjoint(F, tau, gla.colmajor(Wj)))
    ref = O.qr_adjoint(gla.to_numpy(F).astype(np.float64), tau.cpu().numpy().astype(np.float64),
                       np.triu(W).astype(np.float64))
    assert rel(got, ref) <= 50 * max(m, n) * eps_of(dtype), rel(got, ref)


def test_autograd_functions_match_torch_cpu():
    """gla.CholeskyFn / gla.QRFn backward against torch CPU autograd."""
    gla = _gla()
    rng = np.random.default_rng(2)
    n = 96
    A = _spd(n, rng)
    W = np.tril(rng.standard_normal((n, n)))
    Ad = torch.tensor(A, device="cuda", requires_grad=True)
    (gla.CholeskyFn.apply(Ad) * torch.tensor(W, device="cuda")).sum().backward()
    Ac = torch.tensor(A, requires_grad=True)
    (torch.linalg.cholesky(Ac) * torch.tensor(W)).sum().backward()
    g = Ac.grad.numpy()
    assert rel(Ad.grad.cpu().numpy(), 0.5 * (g + g.T)) <= 1e-11

    m, k = 200, 80
    M = rng.standard_normal((m, k))
    Wr = np.triu(rng.standard_normal((k, k)))
    Md = torch.tensor(M, device="cuda", requires_grad=True)
    R = gla.QRFn.apply(Md, 32)
    (R * torch.tensor(Wr, device="cuda")).sum().backward()
    Mc = torch.tensor(M, requires_grad=True)
    _, Rc = torch.linalg.qr(Mc)
    s = np.sign(np.diag(R.detach().cpu().numpy())) * np.sign(np.diag(Rc.detach().numpy()))
    (Rc * torch.tensor(np.diag(s) @ Wr)).sum().backward()
    assert rel(Md.grad.cpu().numpy(), Mc.grad.numpy()) <= 1e-10


def test_adjoint_argument_errors():
    gla = _gla()
    L = gla.colmajor(np.eye(3))
    with pytest.raises(gla.ArgumentError):
        gla.chol_adjoint(L, L, "X")
    F = gla.colmajor(np.ones((3, 5)))
    with pytest.raises(gla.ArgumentError):
        gla.qr_adjoint(F, torch.zeros(3, dtype=torch.float64, device="cuda"),
                       gla.colmajor(np.eye(5)))
