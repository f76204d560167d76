"""Writes the golden fixtures under tests/golden/ from sources independent of
this repository's code: LAPACK through scipy (dpotrf, dgeqrf) and mpmath at
256 bits.  The reference (Julia) cannot run here, so these pin the oracle and
the device path to the mathematics, with Julia's reflector convention applied
where it differs from LAPACK (a 1-element reflector has τ = 2, not 0).

Run:  python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def chol_cases():
    out = {}
    rng = np.random.default_rng(20240101)
    for n in (1, 3, 17, 48):
        B = rng.standard_normal((n, n))
        A = B.T @ B + n * np.eye(n)
        mp.mp.prec = 256
        M = mp.matrix([[mp.mpf(float(A[i, j])) for j in range(n)] for i in range(n)])
        Lm = mp.cholesky(M)
        L = np.array([[float(Lm[i, j]) for j in range(n)] for i in range(n)])
        Ll, info = lapack.dpotrf(A, lower=1)
        assert info == 0 and np.abs(np.tril(Ll) - L).max() <= 1e-12 * np.abs(L).max()
        out[f"A_n{n}"] = A
        out[f"L_n{n}"] = np.tril(L)
    # uniform [0,1) Gram matrix as in test/cholesky.jl:13-17
    rng = np.random.default_rng(123)
    X = rng.random((50, 50))
    A = X.T @ X
    Ll, info = lapack.dpotrf(A, lower=1)
    assert info == 0
    out["A_uniform50"] = A
    out["L_uniform50"] = np.tril(Ll)
    return out


def julia_convention(qr, tau, m, n):
    """LAPACK's dgeqrf output rewritten in Julia's reflector! convention: the
    last reflector of a problem with m ≤ n has length 1, where Julia gives
    τ = 2 and flips the sign of row m of R."""
    qr = qr.copy()
    tau = tau.copy()
    if m <= n:
        qr[m - 1, m - 1:] *= -1
        tau[m - 1] = 2.0
    return qr, tau


def qr_cases():
    out = {}
    rng = np.random.default_rng(777)
    for (m, n) in ((10, 5), (10, 10), (5, 10), (64, 40), (40, 64)):
        A = rng.standard_normal((m, n))
        qr, tau, _, info = lapack.dgeqrf(A)
        assert info == 0
        k = min(m, n)
        F, t = julia_convention(qr, tau[:k], m, n)
        tag = f"{m}x{n}"
        out["A_" + tag] = A
        out["F_" + tag] = F
        out["tau_" + tag] = t
    return out


def main():
    np.savez(os.path.join(HERE, "chol_golden.npz"), **chol_cases())
    np.savez(os.path.join(HERE, "qr_golden.npz"), **qr_cases())
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
