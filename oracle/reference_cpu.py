"""The reference's CPU path (oracle/reference_cpu.c) with its BLAS steps on
the OpenBLAS that numpy bundles, all host threads: cholRecursive! and
qrBlocked! on Float64 as GenericLinearAlgebra.jl runs them (generic loops
single-threaded, BLAS calls threaded).  TEST AND BENCHMARK INFRASTRUCTURE
ONLY: the benchmark's cpu_baseline leg times it and tests/ checks it
against the oracle; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libgla_refcpu.so")

_lib = None
_blas = None


def openblas_path() -> str | None:
    """numpy's bundled ILP64 OpenBLAS (numpy.libs/libscipy_openblas64_*.so)."""
    libs = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), "numpy.libs")
    cands = sorted(glob.glob(os.path.join(libs, "libscipy_openblas64_*.so")))
    return cands[0] if cands else None


def available() -> bool:
    return openblas_path() is not None


def lib() -> C.CDLL:
    """The port with its BLAS hooks installed; raises if either is missing."""
    global _lib, _blas
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    path = openblas_path()
    if path is None:
        raise RuntimeError("numpy's bundled OpenBLAS was not found")
    blas = C.CDLL(path)
    L = C.CDLL(LIB_PATH)
    fns = [C.cast(getattr(blas, name), C.c_void_p)
           for name in ("scipy_cblas_dtrsm64_", "scipy_cblas_dtrmm64_", "scipy_cblas_dgemm64_",
                        "scipy_cblas_ddot64_", "scipy_dtrtri_64_")]
    L.orf_set_blas.argtypes = [C.c_void_p] * 5
    L.orf_set_blas.restype = C.c_int
    if L.orf_set_blas(*fns) != 0:
        raise RuntimeError("OpenBLAS entry points missing")
    L.orf_chol_recursive.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_int64]
    L.orf_chol_recursive.restype = C.c_int64
    L.orf_qr_blocked.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                 C.c_void_p, C.c_void_p, C.c_int64]
    L.orf_qr_blocked.restype = None
    _blas, _lib = blas, L
    return L


def set_threads(n: int) -> int:
    """OpenBLAS threads for the BLAS steps; returns the count in effect."""
    lib()
    _blas.scipy_openblas_set_num_threads64_(C.c_int(int(n)))
    _blas.scipy_openblas_get_num_threads64_.restype = C.c_int
    return int(_blas.scipy_openblas_get_num_threads64_())


def _f(A) -> np.ndarray:
    return np.array(A, dtype=np.float64, order="F", copy=True)


def chol_recursive_inplace(A: np.ndarray, cutoff: int = 1) -> int:
    """cholRecursive!(A, Val{:L}, cutoff) on a Fortran-ordered float64 A in
    place (the benchmark times copy + this); returns info."""
    return int(lib().orf_chol_recursive(A.shape[0], A.ctypes.data, A.shape[0], int(cutoff)))


def chol_recursive(A, cutoff: int = 1):
    """cholRecursive!(copy(A), Val{:L}, cutoff) -> (A_out, info)."""
    A = _f(A)
    info = chol_recursive_inplace(A, cutoff)
    return A, info


def qr_blocked(A, blocksize: int = 12):
    """qrBlocked!(copy(A), blocksize) -> (factors, τ)."""
    A = _f(A)
    m, n = A.shape
    tau = np.zeros(min(m, n))
    work = np.zeros((blocksize, max(n, 1)), order="F")
    lib().orf_qr_blocked(m, n, A.ctypes.data, m, int(blocksize), tau.ctypes.data,
                         work.ctypes.data, blocksize)
    return A, tau
