"""Loader and ctypes signatures of libgla_hip.so (the C ABI in include/gla_hip.h).

The library is required: importing the host API without it raises, so no
call can silently fall back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgla_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "gla_hip.h")

_i64 = C.c_int64
_p = C.c_void_p
_pi64 = C.POINTER(C.c_int64)
_ch = C.c_char
_int = C.c_int

GLA_OK = 0
GLA_ERR_HIP = -1000
GLA_ERR_UNSUPPORTED = -1001
GLA_ERR_TIMEOUT = -1003


def build(jobs: int = 8) -> str:
    """Compile libgla_hip.so for gfx950 with the package Makefile."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)
    return LIB_PATH


# name -> (restype, argtypes); the _d/_s twins share one signature with the
# scalar type substituted.
def _sigs(ft):
    return {
        "gla_chol_unblocked": (_int, [_p, _ch, _i64, _p, _i64, _pi64, _p]),
        "gla_chol_blocked": (_int, [_p, _ch, _i64, _p, _i64, _i64, _pi64, _p]),
        "gla_chol_recursive": (_int, [_p, _ch, _i64, _p, _i64, _i64, _pi64, _p]),
        "gla_rank_update": (_int, [_p, _ch, _i64, _i64, ft, _p, _i64, _p, _i64, _p]),
        "gla_trsm_rlt": (_int, [_p, _i64, _i64, _p, _i64, _p, _i64, _p]),
        "gla_qr_unblocked": (_int, [_p, _i64, _i64, _p, _i64, _p, _p]),
        "gla_qr_blocked": (_int, [_p, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p]),
        "gla_qr_tblock": (_int, [_p, _i64, _i64, _p, _i64, _p, _p, _i64, _p]),
        "gla_wy_apply": (_int, [_p, _ch, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _i64, _p,
                                _i64, _p]),
        "gla_lmul_unit_lower": (_int, [_p, _i64, _i64, _p, _i64, _p, _i64, ft, _p]),
        "gla_gemm": (_int, [_p, _ch, _ch, _i64, _i64, _i64, ft, _p, _i64, _p, _i64, ft, _p,
                            _i64, _p]),
        "gla_chol_solve_batched": (_int, [_p, _p, _p, _p, _p, _p, _p]),
        "gla_chol_adjoint": (_int, [_p, _ch, _i64, _p, _i64, _p, _i64, _p, _i64, _p]),
        "gla_chol_adjoint_batched": (_int, [_p, _ch, _i64, _i64, _p, _i64, _i64, _p, _i64, _i64,
                                            _p, _i64, _i64, _p, _p]),
        "gla_qr_adjoint": (_int, [_p, _i64, _i64, _p, _i64, _p, _p, _i64, _p, _i64, _p]),
        "gla_chol_solve": (_int, [_p, _ch, _i64, _i64, _p, _i64, _p, _i64, _p]),
        "gla_qr_solve": (_int, [_p, _i64, _i64, _i64, _p, _i64, _p, _p, _i64, _p, _p]),
        "gla_ldlt": (_int, [_p, _ch, _i64, _p, _i64, _i64, _p]),
        "gla_symtri": (_int, [_p, _ch, _i64, _p, _i64, _p, _p]),
        "gla_bidiag": (_int, [_p, _i64, _i64, _p, _i64, _p, _p, _p]),
        "gla_ldlt_solve": (_int, [_p, _ch, _i64, _i64, _p, _i64, _p, _i64, _p]),
        "gla_ldlt_lmul": (_int, [_p, _ch, _i64, _i64, _p, _i64, _p, _i64, _p]),
    }


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libgla_hip.so not found at {LIB_PATH}; build it with "
            f"`make -C {PKG_DIR}` (the HIP path has no fallback)")
    L = C.CDLL(LIB_PATH)
    for sfx, ft in (("z", C.c_double), ("c", C.c_float)):
        for name in ("gla_chol_unblocked", "gla_chol_blocked", "gla_chol_recursive",
                     "gla_rank_update", "gla_qr_unblocked", "gla_qr_tblock"):
            f = getattr(L, f"{name}_{sfx}")
            f.restype = _int
            f.argtypes = _sigs(ft)[name][1]
        f = getattr(L, f"gla_qr_blocked_{sfx}")
        f.restype = _int
        f.argtypes = [_p, _i64, _i64, _p, _i64, _i64, _p, _p]
        f = getattr(L, f"gla_wy_apply_{sfx}")
        f.restype = _int
        f.argtypes = [_p, _ch, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _i64, _p]
    for sfx, ft in (("d", C.c_double), ("s", C.c_float)):
        for name, (res, args) in _sigs(ft).items():
            f = getattr(L, f"{name}_{sfx}")
            f.restype = res
            f.argtypes = args
    L.gla_handle_create.argtypes = [C.POINTER(_p)]
    L.gla_handle_create.restype = _int
    L.gla_handle_destroy.argtypes = [_p]
    L.gla_handle_destroy.restype = _int
    L.gla_batch_plan_create.argtypes = [_p, _i64, _p, C.c_int32, C.c_int32, C.POINTER(_p)]
    L.gla_batch_plan_create.restype = _int
    L.gla_batch_plan_destroy.argtypes = [_p]
    L.gla_batch_plan_destroy.restype = _int
    L.gla_chol_solve_mixed_d.argtypes = [_p, _ch, _i64, _p, _i64, _i64, _p, _i64, _p, _i64,
                                         C.c_double, C.c_int32, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_double), _pi64, C.c_char_p, _p]
    L.gla_chol_solve_mixed_d.restype = _int
    L.gla_gemm_lowp_nt.argtypes = [_p, _ch, _i64, _i64, _i64, C.c_float, _p, _i64, _p, _i64, _p,
                                   _i64, _p]
    L.gla_gemm_lowp_nt.restype = _int
    L.gla_chol_recursive_noalloc_d.argtypes = [_ch, _i64, _p, _i64, _i64, _pi64, _p]
    L.gla_chol_recursive_noalloc_d.restype = _int
    L.gla_set_tuning.argtypes = [_p, C.c_char_p, _i64]
    L.gla_set_tuning.restype = _int
    L.gla_get_tuning.argtypes = [_p, C.c_char_p, _pi64]
    L.gla_get_tuning.restype = _int
    L.gla_last_error_string.restype = C.c_char_p
    L.gla_version.restype = C.c_char_p
    _lib = L
    return L


def exported_symbols() -> list[str]:
    """Entry points declared in include/gla_hip.h (parsed from the header)."""
    import re

    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(gla_[a-z0-9_]+)\s*\(", text)))
