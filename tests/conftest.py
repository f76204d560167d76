import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "genericlinearalgebra.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libgla_hip.so")


def rel(x, y) -> float:
    """Normwise relative difference ‖x − y‖_F / ‖y‖_F."""
    cplx = np.iscomplexobj(x) or np.iscomplexobj(y)
    x = np.asarray(x, dtype=np.complex128 if cplx else np.float64)
    y = np.asarray(y, dtype=np.complex128 if cplx else np.float64)
    d = np.linalg.norm(y)
    return float(np.linalg.norm(x - y) / (d if d > 0 else 1.0))


def eps_of(dtype) -> float:
    return float(np.finfo(dtype).eps)


def spd(n, rng, dtype=np.float64):
    """A = BᵀB + n·I with B iid N(0,1): the benchmark's SPD family."""
    B = rng.standard_normal((n, n))
    return (B.T @ B + n * np.eye(n)).astype(dtype)


@pytest.fixture
def rng():
    return np.random.default_rng(20261202)
