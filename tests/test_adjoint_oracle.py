"""Pinning of the oracle's adjoints (oracle.chol_adjoint, oracle.qr_adjoint),
which have no reference code: against PyTorch CPU autograd of
torch.linalg.cholesky / torch.linalg.qr (an independent implementation) and
against central finite differences (h = 1e-6, float64) of the oracle's own
forward factorisations (cholUnblocked!, qrUnblocked!)."""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")


def _spd(n, rng):
    B = rng.standard_normal((n, n))
    return B.T @ B + n * np.eye(n)


@pytest.mark.parametrize("n", [1, 2, 5, 16, 40])
def test_chol_adjoint_vs_torch_autograd(n):
    rng = np.random.default_rng(n)
    A = _spd(n, rng)
    W = np.tril(rng.standard_normal((n, n)))
    At = torch.tensor(A, requires_grad=True)
    L = torch.linalg.cholesky(At)
    (L * torch.tensor(W)).sum().backward()
    g = At.grad.numpy()
    gs = 0.5 * (g + g.T)                 # symmetric part, whatever torch's convention
    Lo, info = O.chol_unblocked(A)
    assert info == 0
    ours = O.chol_adjoint(np.tril(Lo), W, "S")
    assert np.allclose(ours, ours.T, rtol=0, atol=1e-14 * np.abs(ours).max())
    assert np.linalg.norm(ours - gs) <= 1e-10 * max(1.0, np.linalg.norm(gs))


def test_chol_adjoint_lower_convention_vs_central_fd():
    """d/dA_ij of <W, chol(A)> over the lower triangle the factorisation reads."""
    n, h = 7, 1e-6
    rng = np.random.default_rng(3)
    A = _spd(n, rng)
    W = np.tril(rng.standard_normal((n, n)))

    def f(X):
        Lx, _ = O.chol_unblocked(X)
        return float(np.sum(np.tril(Lx) * W))

    fd = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            E = np.zeros((n, n))
            E[i, j] = h
            fd[i, j] = (f(A + E) - f(A - E)) / (2 * h)
    Lo, _ = O.chol_unblocked(A)
    ours = O.chol_adjoint(np.tril(Lo), W, "L")
    assert np.allclose(np.triu(ours, 1), 0)
    assert np.linalg.norm(ours - fd) <= 1e-6 * np.linalg.norm(fd)
    # a plausible slip — forgetting to halve the diagonal in Φ — is caught
    P = np.tril(np.tril(Lo).T @ W)
    wrong = np.linalg.solve(np.tril(Lo).T, np.linalg.solve(np.tril(Lo).T, 0.5 * (P + P.T)).T).T
    wrong = np.tril(2 * wrong) - np.diag(np.diag(wrong))
    assert np.linalg.norm(wrong - fd) > 1e-3 * np.linalg.norm(fd)


@pytest.mark.parametrize("m,n", [(1, 1), (6, 6), (10, 4), (40, 17)])
def test_qr_adjoint_vs_torch_autograd(m, n):
    rng = np.random.default_rng(m * 100 + n)
    A = rng.standard_normal((m, n))
    W = np.triu(rng.standard_normal((n, n)))
    F, tau = O.qr_unblocked(A)
    R = np.triu(F[:n, :n])
    At = torch.tensor(A, requires_grad=True)
    Qt, Rt = torch.linalg.qr(At, mode="reduced")
    # same factorisation up to the signs of R's rows; flip W's rows to match
    s = np.sign(np.diag(R)) * np.sign(np.diag(Rt.detach().numpy()))
    assert np.allclose(np.diag(s) @ Rt.detach().numpy(), R, atol=1e-12 * np.abs(R).max())
    (Rt * torch.tensor(np.diag(s) @ W)).sum().backward()
    ours = O.qr_adjoint(F, tau, W)
    g = At.grad.numpy()
    assert np.linalg.norm(ours - g) <= 1e-10 * max(1.0, np.linalg.norm(g))


def test_qr_adjoint_vs_central_fd():
    m, n, h = 9, 5, 1e-6
    rng = np.random.default_rng(8)
    A = rng.standard_normal((m, n))
    W = np.triu(rng.standard_normal((n, n)))

    def f(X):
        Fx, _ = O.qr_unblocked(X)
        return float(np.sum(np.triu(Fx[:n, :n]) * W))

    fd = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            E = np.zeros((m, n))
            E[i, j] = h
            fd[i, j] = (f(A + E) - f(A - E)) / (2 * h)
    F, tau = O.qr_unblocked(A)
    ours = O.qr_adjoint(F, tau, W)
    assert np.linalg.norm(ours - fd) <= 1e-6 * np.linalg.norm(fd)


def test_thin_q_is_orthonormal_and_reproduces_a():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((12, 5))
    F, tau = O.qr_unblocked(A)
    Q = O.thin_q(F, tau)
    assert np.allclose(Q.T @ Q, np.eye(5), atol=1e-14)
    assert np.allclose(Q @ np.triu(F[:5, :5]), A, atol=1e-13)
