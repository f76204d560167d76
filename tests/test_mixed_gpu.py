"""Mixed-precision Cholesky solve with Float64 iterative refinement, and the
low-precision MFMA product it is built on.

The product is checked against a float32 torch product of the same
bf16/fp16-rounded operands (exact products, fp32 sums: tolerance
k·eps32 relative to |A||B|ᵀ).  The solver is checked by what it promises:
the backward error it reports is recomputed here in float64 on the CPU and
must be ≤ 1e-12, and X must agree with a float64 LAPACK solve to within
κ·1e-12.  Matrices: A = Q diag(λ) Qᵀ with λ log-spaced over [1/κ, 1]."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def spd_kappa(n, kappa, rng):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(0.0, -np.log10(kappa), n)
    A = (Q * lam) @ Q.T
    return (A + A.T) / 2


def backward_error(A, X, B):
    R = B - A @ X
    an = np.linalg.norm(A, "fro")
    return max(np.linalg.norm(R[:, k]) / (an * np.linalg.norm(X[:, k]) + np.linalg.norm(B[:, k]))
               for k in range(B.shape[1]))


@pytest.mark.parametrize("prec,tdt", [("bf16x1", torch.bfloat16), ("fp16x1", torch.float16)])
@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (300, 200, 77), (517, 129, 256), (64, 1000, 13)])
@pytest.mark.parametrize("alpha", [1.0, -1.0])
def test_lowp_gemm_vs_torch(prec, tdt, m, n, k, alpha):
    gla = _gla()
    g = torch.Generator().manual_seed(m + n + k)
    A = torch.randn(m, k, generator=g)
    B = torch.randn(n, k, generator=g)
    C0 = torch.randn(m, n, generator=g)
    Ar = A.to(tdt).float()
    Br = B.to(tdt).float()
    ref = C0.double() + alpha * (Ar.double() @ Br.double().T)
    Cd = gla.colmajor(C0, torch.float32)
    gla.gemm_lowp_nt_(Cd, gla.colmajor(A, torch.float32), gla.colmajor(B, torch.float32),
                      alpha=alpha, precision=prec)
    out = gla.to_numpy(Cd).astype(np.float64)
    scale = np.abs(C0.double().numpy()) + np.abs(Ar.double().numpy()) @ np.abs(Br.double().numpy()).T
    err = np.abs(out - ref.numpy()) / scale
    assert err.max() <= (k + 2) * np.finfo(np.float32).eps, err.max()


def _split(X, tdt, nterm, scale):
    """The terms the device splits X·scale into: each the rounding to tdt of
    what the previous ones leave (float64 here, exact for these values)."""
    x = X.double() * scale
    out = []
    for _ in range(nterm):
        t = x.to(tdt).double()
        out.append(t)
        x = x - t
    return out


@pytest.mark.parametrize("prec,tdt,nterm", [("bf16", torch.bfloat16, 3),
                                            ("fp16", torch.float16, 2)])
@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (300, 200, 77), (517, 129, 256)])
def test_split_gemm_vs_terms(prec, tdt, nterm, m, n, k):
    """The multi-term product: A'B'ᵀ over the segments [A₀A₀A₁]·[B₀B₁B₀]ᵀ
    (FP16, after a common power-of-two scale s = 2^(14−e), max < 2^e) or
    the six BF16 segments, against float64 sums of the same products of the
    exact terms; tolerance K'·eps32 relative to the summed magnitudes, and
    against the float64 product itself at the split's accuracy (2⁻²⁰)."""
    gla = _gla()
    g = torch.Generator().manual_seed(7 * m + n + k)
    A = torch.randn(m, k, generator=g) * 1e-3
    B = torch.randn(n, k, generator=g) * 1e-3
    C0 = torch.randn(m, n, generator=g) * 1e-6
    mx = max(float(A.abs().max()), float(B.abs().max()))
    sc = 2.0 ** (14 - np.frexp(mx)[1]) if prec == "fp16" else 1.0
    ta, tb = _split(A, tdt, nterm, sc), _split(B, tdt, nterm, sc)
    pairs = [(0, 0), (0, 1), (1, 0)] + ([(0, 2), (2, 0), (1, 1)] if nterm == 3 else [])
    ref = C0.double() - sum(ta[i] @ tb[j].T for i, j in pairs) / (sc * sc)
    mag = np.abs(C0.double().numpy()) + sum(
        np.abs(ta[i].numpy()) @ np.abs(tb[j].numpy()).T for i, j in pairs) / (sc * sc)
    Cd = gla.colmajor(C0, torch.float32)
    gla.gemm_lowp_nt_(Cd, gla.colmajor(A, torch.float32), gla.colmajor(B, torch.float32),
                      alpha=-1.0, precision=prec)
    out = gla.to_numpy(Cd).astype(np.float64)
    kk = k * len(pairs) + 2
    assert (np.abs(out - ref.numpy()) / mag).max() <= kk * np.finfo(np.float32).eps
    exact = C0.double() - A.double() @ B.double().T
    assert (np.abs(out - exact.numpy()) / mag).max() <= 2.0 ** -20 + kk * np.finfo(np.float32).eps


def _check_solution(A, B, X, res, kappa, tol=1e-12):
    be = backward_error(A, X, B)
    assert res.converged and res.berr <= tol, res
    assert be <= 2 * tol, (be, res)
    assert abs(be - res.berr) <= 0.5 * max(be, res.berr) + 1e-16
    Xr = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= kappa * 1e-11


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_auto_precision_reaches_tolerance(kappa):
    gla = _gla()
    rng = np.random.default_rng(int(np.log10(kappa)))
    n, nrhs = 700, 16
    A = spd_kappa(n, kappa, rng)
    B = rng.standard_normal((n, nrhs))
    Ad, Bd = gla.colmajor(A), gla.colmajor(B)
    X, res = gla.chol_solve_mixed(Ad, Bd, precision="auto", tol=1e-12, max_iter=60)
    _check_solution(A, B, gla.to_numpy(X), res, kappa)
    # A and B are inputs only
    assert np.array_equal(gla.to_numpy(Ad), A)
    assert np.array_equal(gla.to_numpy(Bd), B)


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32", "fp64"])
@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_every_precision_reaches_tolerance(prec, kappa):
    """The benchmark's matrix family at n = 2048 with 16 right-hand sides:
    every precision, the BF16 and FP16 multi-term factorisations included,
    reaches η ≤ 1e-12 at κ = 1e2, 1e4 and 1e6, and the returned X matches a
    float64 LAPACK solve to κ·1e-11."""
    gla = _gla()
    rng = np.random.default_rng(int(np.log10(kappa)) + 10)
    n, nrhs = 2048, 16
    A = spd_kappa(n, kappa, rng)
    B = rng.standard_normal((n, nrhs))
    X, res = gla.chol_solve_mixed(gla.colmajor(A), gla.colmajor(B), precision=prec,
                                  tol=1e-12, max_iter=30)
    assert res.precision == prec
    _check_solution(A, B, gla.to_numpy(X), res, kappa)


@pytest.mark.parametrize("prec", ["bf16x1", "fp16x1"])
@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_single_term_converges_or_says_so(prec, kappa):
    """One BF16/FP16 term: converges at κ = 1e2; at κ = 1e6 it either
    raises DomainError (the factorisation breaks down) or reports, truthfully,
    that it did not converge."""
    gla = _gla()
    rng = np.random.default_rng(int(np.log10(kappa)) + 20)
    n, nrhs = 1024, 16
    A = spd_kappa(n, kappa, rng)
    B = rng.standard_normal((n, nrhs))
    try:
        X, res = gla.chol_solve_mixed(gla.colmajor(A), gla.colmajor(B), precision=prec,
                                      tol=1e-12, max_iter=30)
    except gla.DomainError:
        assert kappa > 1e2
        return
    assert res.precision == prec
    Xh = gla.to_numpy(X)
    if kappa <= 1e2 or res.converged:
        _check_solution(A, B, Xh, res, kappa)
    else:
        be = backward_error(A, Xh, B)
        assert not (be <= 1e-12) and not (res.berr <= 1e-12)


def test_more_than_sixteen_rhs_and_lower_triangle_only():
    """37 right-hand sides (three passes of 16); the strict upper triangle of
    A holds garbage that must not be read."""
    gla = _gla()
    rng = np.random.default_rng(21)
    n = 300
    A = spd_kappa(n, 1e3, rng)
    B = rng.standard_normal((n, 37))
    Ag = np.tril(A) + np.triu(np.full((n, n), 1e6), 1)
    X, res = gla.chol_solve_mixed(gla.colmajor(Ag), gla.colmajor(B), precision="auto")
    _check_solution(A, B, gla.to_numpy(X), res, 1e3)


def test_fp64_needs_no_correction_on_wellconditioned():
    gla = _gla()
    rng = np.random.default_rng(3)
    A = spd_kappa(300, 10.0, rng)
    B = rng.standard_normal((300, 4))
    X, res = gla.chol_solve_mixed(gla.colmajor(A), gla.colmajor(B), precision="fp64",
                                  tol=1e-12)
    assert res.iters == 0 and res.berr <= 1e-12 and res.precision == "fp64"


def test_ragged_sizes_and_vector_rhs():
    gla = _gla()
    rng = np.random.default_rng(9)
    for n in (1, 5, 127, 129, 383):
        A = spd_kappa(n, 100.0, rng)
        b = rng.standard_normal(n)
        x, res = gla.chol_solve_mixed(gla.colmajor(A), torch.from_numpy(b).cuda(),
                                      precision="bf16", tol=1e-12)
        xh = x.cpu().numpy()
        assert xh.shape == (n,)
        assert res.converged, (n, res)
        assert backward_error(A, xh[:, None], b[:, None]) <= 2e-12, (n, res)


def test_not_spd_raises_domain_error():
    gla = _gla()
    rng = np.random.default_rng(4)
    A = spd_kappa(256, 10.0, rng)
    A[200, 200] = -5.0
    with pytest.raises(gla.DomainError):
        gla.chol_solve_mixed(gla.colmajor(A), gla.colmajor(rng.standard_normal((256, 2))),
                             precision="bf16")


def test_bad_precision_rejected():
    gla = _gla()
    A = gla.colmajor(np.eye(4))
    with pytest.raises(gla.ArgumentError):
        gla.chol_solve_mixed(A, gla.colmajor(np.ones((4, 1))), precision="fp8")
