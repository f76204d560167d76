"""The no-allocation Cholesky entry point, gla_chol_recursive_noalloc_d,
against gla_chol_recursive_d (cholRecursive!, src/cholesky.jl:37-55) through
the C ABI.  Calls follow one another on one GPU, the second smaller than the
first, so the fixed scratch array still holds the first call's contents:
each result must equal the handle entry point's recursive-halving driver
("chol_block" 0) bit for bit and its default driver to 10·n·eps, with the
same info."""
import ctypes as C

import numpy as np
import pytest

from conftest import eps_of, rel, spd

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _noalloc(A, uplo="L", cutoff=1):
    gla = _gla()
    info = C.c_int64(-1)
    rc = gla.lib().gla_chol_recursive_noalloc_d(uplo.encode(), A.shape[0], C.c_void_p(A.data_ptr()),
                                                A.stride(1), cutoff, C.byref(info), _stream())
    assert rc == 0, gla.lib().gla_last_error_string()
    return info.value


def _handle_entry(A, uplo="L", cutoff=1, chol_block=-1):
    gla = _gla()
    info = C.c_int64(-1)
    with gla.tuning(chol_block=chol_block):
        rc = gla.lib().gla_chol_recursive_d(gla._handle(0), uplo.encode(), A.shape[0],
                                            C.c_void_p(A.data_ptr()), A.stride(1), cutoff,
                                            C.byref(info), _stream())
    assert rc == 0, gla.lib().gla_last_error_string()
    return info.value


def _check_pair(A0, uplo):
    gla = _gla()
    n = A0.shape[0]
    X = gla.colmajor(A0)
    Y = gla.colmajor(A0)
    Z = gla.colmajor(A0)
    i_x = _noalloc(X, uplo)
    i_y = _handle_entry(Y, uplo, chol_block=0)
    i_z = _handle_entry(Z, uplo)
    torch.cuda.synchronize()
    assert i_x == i_y == i_z
    if i_x == 0:                  # after a negative pivot the contents are NaN-laden
        assert torch.equal(X, Y), "differs from the recursive-halving driver"
        tri = torch.tril if uplo == "L" else torch.triu
        e = rel(tri(X).cpu().numpy(), tri(Z).cpu().numpy())
        assert e <= 10 * n * eps_of(np.float64), e
    return i_x


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_noalloc_matches_handle_entry_on_shrinking_calls(uplo):
    rng = np.random.default_rng(91)
    for n in (3000, 1100, 130):          # each call smaller than the one before
        assert _check_pair(spd(n, rng), uplo) == 0


def test_noalloc_info_after_a_larger_call():
    rng = np.random.default_rng(92)
    assert _check_pair(spd(2600, rng), "L") == 0
    A = spd(900, rng)
    A[500, 500] = -1e6                    # the first negative pivot is 501
    assert _check_pair(A, "L") == 501
    assert _check_pair(spd(40, rng), "L") == 0   # info cleared by the next call


def test_noalloc_rejects_sizes_beyond_its_scratch():
    gla = _gla()
    A = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = gla.lib().gla_chol_recursive_noalloc_d(b"L", 32769, C.c_void_p(A.data_ptr()), 32769, 1,
                                                None, _stream())
    assert rc == gla._lib.GLA_ERR_UNSUPPORTED
    assert b"maximum" in gla.lib().gla_last_error_string()
