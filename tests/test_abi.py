"""The C-ABI library loads and exports every entry point include/gla_hip.h
declares; argument errors are reported without touching a device.  CPU only:
no compute call is made."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gla_hip.h")


def _lib():
    from gla import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgla_hip.so not built (run __graft_entry__.build())")
    return _lib


def test_header_declares_reference_entry_points():
    text = open(HEADER).read()
    names = set(re.findall(r"\b(gla_[a-z0-9_]+)\s*\(", text))
    for base in ("chol_unblocked", "chol_blocked", "chol_recursive", "rank_update", "trsm_rlt",
                 "qr_unblocked", "qr_blocked", "qr_tblock", "wy_apply", "lmul_unit_lower", "gemm"):
        for sfx in ("d", "s"):
            assert f"gla_{base}_{sfx}" in names
    assert "src/cholesky.jl" in text and "src/qr.jl" in text and "src/householder.jl" in text


def _declarations_with_comments(text):
    """(name, the comment block nearest above its declaration) for every
    function the header declares."""
    out = []
    for m in re.finditer(r"^(?:int|const char \*)\s*(gla_[a-z0-9_]+)\s*\(", text, re.M):
        head = text[:m.start()]
        end = head.rfind("*/")
        start = head.rfind("/*", 0, end)
        out.append((m.group(1), head[start:end] if end >= 0 and start >= 0 else ""))
    return out


def test_every_declaration_cites_the_reference_or_says_it_has_none():
    text = open(HEADER).read()
    decls = _declarations_with_comments(text)
    assert len(decls) >= 40
    cite = re.compile(r"src/[a-zA-Z]+\.jl:\d+|[Nn]o (reference )?counterpart")
    bad = [name for name, comment in decls if not cite.search(comment)]
    assert not bad, f"declarations without a reference citation: {bad}"


def test_library_exports_every_declared_symbol():
    L = _lib()
    lib = C.CDLL(L.LIB_PATH)
    missing = [s for s in L.exported_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string_need_no_device():
    L = _lib()
    lib = L.lib()
    assert b"gfx950" in lib.gla_version()
    assert isinstance(lib.gla_last_error_string(), bytes)


def test_null_handle_is_an_argument_error():
    L = _lib()
    lib = L.lib()
    rc = lib.gla_chol_recursive_d(None, b"L", 4, None, 4, 1, None, None)
    assert rc == -1
    rc = lib.gla_qr_blocked_d(None, 4, 4, None, 4, 2, None, None, 2, None)
    assert rc == -1


def test_noalloc_entry_rejects_arguments_before_touching_a_device():
    """gla_chol_recursive_noalloc_d checks its arguments first: a size past
    its fixed scratch and a bad uplo return at once (no GPU here)."""
    L = _lib()
    lib = L.lib()
    rc = lib.gla_chol_recursive_noalloc_d(b"L", 32769, None, 32769, 1, None, None)
    assert rc == L.GLA_ERR_UNSUPPORTED
    assert b"32768" in lib.gla_last_error_string()
    assert lib.gla_chol_recursive_noalloc_d(b"X", 4, None, 4, 1, None, None) == -1
    assert lib.gla_chol_recursive_noalloc_d(b"L", -1, None, 4, 1, None, None) == -2


def test_host_api_fails_loudly_without_library(monkeypatch, tmp_path):
    from gla import _lib as L

    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(RuntimeError, match="no fallback"):
        L.lib()
