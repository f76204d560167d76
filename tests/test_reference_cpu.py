"""The reference's CPU path used as the benchmark's CPU baseline
(oracle/reference_cpu.c: generic loops + OpenBLAS for the BLAS steps)
computes what the oracle computes: cholRecursive! and qrBlocked! agree to
10·n·eps, the DomainError index agrees, and the QR agrees on the reference's
own shape × block-size grid (test/qr.jl:7-19).  CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel

R = pytest.importorskip("reference_cpu")

pytestmark = pytest.mark.skipif(not R.available(), reason="numpy's bundled OpenBLAS not found")


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 255, 600])
@pytest.mark.parametrize("cutoff", [1, 4, 50])
def test_chol_recursive_matches_oracle(rng, n, cutoff):
    B = rng.standard_normal((n, n))
    A = B.T @ B + n * np.eye(n)
    L, info = R.chol_recursive(A, cutoff)
    Lo, io = O.chol_recursive(A, cutoff)
    assert info == io == 0
    assert rel(np.tril(L), np.tril(Lo)) <= 10 * n * eps_of(np.float64)
    assert np.array_equal(np.triu(L, 1), np.triu(A, 1))


def test_chol_domain_error_index(rng):
    n, k = 90, 40
    B = rng.standard_normal((n, n))
    A = B.T @ B + n * np.eye(n)
    Lr = np.linalg.cholesky(A)
    A[k, k] -= Lr[k, k] ** 2 + 1.0
    _, info = R.chol_recursive(A)
    _, io = O.chol_recursive(A)
    assert info == io == k + 1


@pytest.mark.parametrize("m,n", [(10, 5), (10, 10), (5, 10), (100, 50), (100, 100), (50, 100),
                                 (700, 300)])
@pytest.mark.parametrize("bz", [1, 3, 8, 17, 33, 128])
def test_qr_blocked_matches_oracle(rng, m, n, bz):
    A = rng.standard_normal((m, n))
    F, tau = R.qr_blocked(A, bz)
    Fo, to = O.qr_blocked(A, bz)
    tol = 10 * max(m, n) * eps_of(np.float64)
    assert rel(F, Fo) <= tol
    assert rel(tau, to) <= tol


def test_threads_can_be_set():
    assert R.set_threads(2) == 2
    assert R.set_threads(4) == 4
