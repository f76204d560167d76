"""MFMA GEMM engine (the contraction under SYRK / TRSM / WY update) against
a float64 numpy product on the same inputs."""
import numpy as np
import pytest

from conftest import eps_of, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (37, 53, 29), (128, 128, 128),
                                   (200, 300, 17), (513, 257, 1031), (64, 2048, 4096)])
def test_gemm_matches_numpy(rng, dtype, ta, tb, m, n, k):
    gla = _gla()
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C0 = rng.standard_normal((m, n))
    alpha, beta = -0.75, 0.5
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    Ad, Bd, Cd = (A.astype(dtype), B.astype(dtype), C0.astype(dtype))
    ref = (alpha * (opA.astype(dtype).astype(np.float64) @ opB.astype(dtype).astype(np.float64))
           + beta * Cd.astype(np.float64))
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    Cg = gla.colmajor(Cd, tdt)
    gla.gemm_(ta, tb, alpha, gla.colmajor(Ad, tdt), gla.colmajor(Bd, tdt), beta, Cg)
    out = gla.to_numpy(Cg)
    # elementwise bound: |error| <= c·k·eps·(|α||A||B| + |β||C|)
    bound = (abs(alpha) * (np.abs(opA) @ np.abs(opB)) + abs(beta) * np.abs(C0)) * k * eps_of(dtype) * 4
    assert np.all(np.abs(out - ref) <= bound + 1e-300), rel(out, ref)


def test_gemm_unaligned_submatrix(rng):
    """Odd leading dimension and an odd row offset force the scalar loads."""
    gla = _gla()
    big = rng.standard_normal((71, 90))
    Ag = gla.colmajor(big)
    sub = Ag[3:70, 5:45]            # 67×40, ld 71, base not 16-byte aligned
    Bm = rng.standard_normal((40, 33))
    Cm = np.zeros((67, 33))
    Cg = gla.colmajor(Cm)
    gla.gemm_("N", "N", 1.0, sub, gla.colmajor(Bm), 0.0, Cg)
    ref = big[3:70, 5:45] @ Bm
    assert rel(gla.to_numpy(Cg), ref) < 64 * 40 * eps_of(np.float64)


VARIANTS = [("gemm_dl", 1), ("gemm_dl", 2), ("gemm_dl", 3), ("gemm_lds1", 0)]
# each variant is compared with the register-staged, single-LDS-buffer kernel
BASE = {"gemm_dl": 0, "gemm_lds1": 1}


@pytest.mark.parametrize("knob,val", VARIANTS)
@pytest.mark.parametrize("m,n,k", [(1024, 768, 512), (1000, 900, 256), (129, 383, 16),
                                   (2048, 2048, 1024)])
def test_kernel_variants_match_default(rng, knob, val, m, n, k):
    """The FP64 GEMM variants (direct-to-LDS ring; double-buffered LDS) on
    NT products with 128×128 tiles, ragged M/N edges included, give the
    default kernel's result: same MFMA order, so the same bits."""
    gla = _gla()
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    C0 = rng.standard_normal((m, n))
    outs = []
    for v in (BASE[knob], val):
        Cg = gla.colmajor(C0)
        with gla.tuning(gemm_dl=0, **{knob: v}) if knob != "gemm_dl" else gla.tuning(gemm_dl=v):
            gla.gemm_("N", "T", -1.0, gla.colmajor(A), gla.colmajor(B), 1.0, Cg)
        outs.append(gla.to_numpy(Cg))
    ref = C0 - A @ B.T
    assert rel(outs[1], ref) <= 8 * k * eps_of(np.float64)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("knob,val", VARIANTS)
def test_kernel_variants_syrk_lower(rng, knob, val):
    gla = _gla()
    n, k = 1536, 512
    X = rng.standard_normal((n, k))
    C0 = rng.standard_normal((n, n))
    outs = []
    for v in (BASE[knob], val):
        Cg = gla.colmajor(C0)
        with gla.tuning(gemm_dl=0, **{knob: v}) if knob != "gemm_dl" else gla.tuning(gemm_dl=v):
            gla.rankUpdate_(gla.Hermitian(Cg, "L"), gla.colmajor(X), -1.0)
        outs.append(gla.to_numpy(Cg))
    ref = C0 - X @ X.T
    assert rel(np.tril(outs[1]), np.tril(ref)) <= 8 * k * eps_of(np.float64)
    assert np.array_equal(np.triu(outs[1], 1), np.triu(C0, 1))
    assert np.array_equal(outs[0], outs[1])
