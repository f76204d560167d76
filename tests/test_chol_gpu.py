"""Cholesky on the device against the oracle (CPU restatement of
src/cholesky.jl), closed-form answers and the reference's own test cases
(test/cholesky.jl:4-30).  Bar: normwise relative difference of the factor
≤ 10·n·eps of the element type."""
import numpy as np
import pytest

import oracle as O
from conftest import eps_of, rel, spd

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gla():
    import gla

    return gla


TDT = {np.float64: "float64", np.float32: "float32"}


def _tdt(dtype):
    return getattr(torch, TDT[dtype])


def _run(fn, A_np, dtype, *args):
    gla = _gla()
    Ag = gla.colmajor(A_np.astype(dtype), _tdt(dtype))
    out = fn(Ag, *args)
    torch.cuda.synchronize()
    return gla.to_numpy(Ag), out


VARIANTS = [
    ("unblocked", lambda g: (lambda A: g.cholUnblocked_(A, g.Val.L)), lambda A: O.chol_unblocked(A)),
    ("blocked5", lambda g: (lambda A: g.cholBlocked_(A, g.Val.L, 5)), lambda A: O.chol_blocked(A, 5)),
    ("blocked10", lambda g: (lambda A: g.cholBlocked_(A, g.Val.L, 10)), lambda A: O.chol_blocked(A, 10)),
    ("recursive1", lambda g: (lambda A: g.cholRecursive_(A, g.Val.L, 1)), lambda A: O.chol_recursive(A, 1)),
    ("recursive4", lambda g: (lambda A: g.cholRecursive_(A, g.Val.L, 4)), lambda A: O.chol_recursive(A, 4)),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name,dev,orc", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_reference_case_n50(dtype, name, dev, orc):
    """test/cholesky.jl:8-27: n = 50, A = rand(n,n), AcA = A'A."""
    rng = np.random.default_rng(123)
    A = rng.random((50, 50)).astype(dtype)
    AcA = (A.T @ A).astype(dtype)
    out, _ = _run(dev(_gla()), AcA, dtype)
    ref, info = orc(AcA)
    assert info == 0
    n = 50
    assert rel(np.tril(out), np.tril(ref)) <= 10 * n * eps_of(dtype)
    # the strict upper triangle keeps the input (the reference never writes it)
    assert np.array_equal(np.triu(out, 1), np.triu(AcA, 1))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 2, 3, 64, 127, 128, 129, 255, 256, 300, 777, 1000])
@pytest.mark.parametrize("algo", ["unblocked", "blocked", "recursive"])
def test_sizes_against_oracle(rng, dtype, n, algo):
    gla = _gla()
    A = spd(n, rng, dtype)
    if algo == "unblocked":
        out, _ = _run(lambda X: gla.cholUnblocked_(X, gla.Val.L), A, dtype)
        ref, info = O.chol_unblocked(A)
    elif algo == "blocked":
        out, _ = _run(lambda X: gla.cholBlocked_(X, gla.Val.L, 64), A, dtype)
        ref, info = O.chol_blocked(A, 64)
    else:
        out, ret = _run(lambda X: gla.cholRecursive_(X, gla.Val.L, 1), A, dtype)
        assert isinstance(ret, gla.LowerTriangular)
        ref, info = O.chol_recursive(A, 1)
    assert info == 0
    assert rel(np.tril(out), np.tril(ref)) <= 10 * n * eps_of(dtype)
    assert np.array_equal(np.triu(out, 1), np.triu(A, 1))


@pytest.mark.parametrize("n", [2500, 4100])
@pytest.mark.parametrize("block", [-1, 0, 256, 512, 768, 1024])
def test_large_recursive_backward_error(rng, n, block):
    """Both device drivers behind cholRecursive!: recursive halving past the
    2048 leaf (block 0) and the wide-block look-ahead (block 256 to 1024);
    checked against LAPACK's factor and by ‖A − LLᵀ‖/‖A‖."""
    gla = _gla()
    A = spd(n, rng)
    with gla.tuning(chol_block=block):
        out, _ = _run(lambda X: gla.cholRecursive_(X, gla.Val.L), A, np.float64)
    L = np.tril(out)
    Lr = np.linalg.cholesky(A)
    assert rel(L, Lr) <= 10 * n * eps_of(np.float64)
    assert rel(L @ L.T, A) <= 10 * n * eps_of(np.float64)


def test_closed_form_min_matrix():
    """A_ij = min(i,j) (1-based) has L = tril(ones) exactly."""
    gla = _gla()
    n = 200
    i = np.arange(1, n + 1)
    A = np.minimum.outer(i, i).astype(np.float64)
    out, _ = _run(lambda X: gla.cholRecursive_(X, gla.Val.L), A, np.float64)
    assert np.array_equal(np.tril(out), np.tril(np.ones((n, n))))


def test_closed_form_pascal():
    """Symmetric Pascal S_ij = C(i+j, i) (0-based) has L_ij = C(i, j), exact for n ≤ 28."""
    from math import comb

    gla = _gla()
    n = 28
    S = np.array([[comb(i + j, i) for j in range(n)] for i in range(n)], dtype=np.float64)
    L = np.array([[comb(i, j) if j <= i else 0 for j in range(n)] for i in range(n)], dtype=np.float64)
    out, _ = _run(lambda X: gla.cholBlocked_(X, gla.Val.L, 10), S, np.float64)
    assert rel(np.tril(out), L) <= 10 * n * eps_of(np.float64)


def test_closed_form_2x2():
    gla = _gla()
    out, _ = _run(lambda X: gla.cholUnblocked_(X, gla.Val.L), np.array([[4.0, 2.0], [2.0, 3.0]]),
                  np.float64)
    assert np.allclose(np.tril(out), [[2.0, 0.0], [1.0, np.sqrt(2.0)]], rtol=0, atol=4e-16)


@pytest.mark.parametrize("k", [0, 5, 130, 299])
def test_domain_error_index(rng, k):
    """A negative pivot at position k+1 raises DomainError naming it
    (the reference's sqrt throws at src/cholesky.jl:5,40)."""
    gla = _gla()
    n = 300
    A = spd(n, rng)
    # make the (k+1)-th pivot negative: subtract enough from A[k,k]
    L = np.linalg.cholesky(A)
    A[k, k] -= L[k, k] ** 2 + 1.0
    _, info = O.chol_recursive(A)
    assert info == k + 1
    for fn in (lambda X: gla.cholUnblocked_(X, gla.Val.L),
               lambda X: gla.cholBlocked_(X, gla.Val.L, 10),
               lambda X: gla.cholRecursive_(X, gla.Val.L)):
        with pytest.raises(gla.DomainError, match=str(k + 1)):
            _run(fn, A, np.float64)


def test_zero_pivot_is_not_an_error():
    """sqrt(0) does not throw in the reference; Inf/NaN propagate silently."""
    gla = _gla()
    A = np.array([[0.0, 0.0], [0.0, 1.0]])
    out, _ = _run(lambda X: gla.cholUnblocked_(X, gla.Val.L), A, np.float64)
    ref, info = O.chol_unblocked(A)
    assert info == 0
    assert out[0, 0] == 0.0


@pytest.mark.parametrize("n", [50, 300])
def test_upper_extension(rng, n):
    """Val{:U} (no reference counterpart): U = Lᵀ in the upper triangle,
    the strict lower triangle keeps its input."""
    gla = _gla()
    A = spd(n, rng)
    X = np.triu(A) + np.tril(rng.standard_normal((n, n)), -1)   # garbage below
    out, ret = _run(lambda Y: gla.cholRecursive_(Y, gla.Val.U), X, np.float64)
    assert isinstance(ret, gla.UpperTriangular)
    Lr = np.linalg.cholesky(A)
    assert rel(np.triu(out), Lr.T) <= 10 * n * eps_of(np.float64)
    assert np.array_equal(np.tril(out, -1), np.tril(X, -1))


def test_non_square_dimension_mismatch():
    gla = _gla()
    with pytest.raises(gla.DimensionMismatch):
        gla.cholUnblocked_(gla.colmajor(np.ones((3, 4))), gla.Val.L)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rank_update_generic(rng, dtype):
    """rankUpdate!(Hermitian(C,:L), A, α) and (…, :U) touch only their
    triangle (src/juliaBLAS.jl:89-112); rank-1 form with a vector."""
    gla = _gla()
    n, k = 77, 13
    C0 = rng.standard_normal((n, n)).astype(dtype)
    A = rng.standard_normal((n, k)).astype(dtype)
    for uplo in ("L", "U"):
        Cg = gla.colmajor(C0, _tdt(dtype))
        gla.rankUpdate_(gla.Hermitian(Cg, uplo), gla.colmajor(A, _tdt(dtype)), -0.5)
        out = gla.to_numpy(Cg)
        full = C0.astype(np.float64) - 0.5 * A.astype(np.float64) @ A.astype(np.float64).T
        tri = np.tril if uplo == "L" else np.triu
        other = (lambda X: np.triu(X, 1)) if uplo == "L" else (lambda X: np.tril(X, -1))
        assert rel(tri(out), tri(full)) <= 10 * n * eps_of(dtype)
        assert np.array_equal(other(out), other(C0))
    x = rng.standard_normal(n).astype(dtype)
    Cg = gla.colmajor(C0, _tdt(dtype))
    gla.rankUpdate_(gla.Hermitian(Cg, "L"), gla.colmajor(x, _tdt(dtype)), 2.0)
    ref = O.syr_l(C0, x, 2.0)
    assert rel(np.tril(gla.to_numpy(Cg)), np.tril(ref)) <= 10 * n * eps_of(dtype)


@pytest.mark.parametrize("m,n", [(1, 1), (50, 7), (300, 128), (257, 300), (1000, 700)])
def test_rdiv_lower_adjoint(rng, m, n):
    """rdiv!(B, LowerTriangular(L)') as at src/cholesky.jl:28,48."""
    gla = _gla()
    L = np.linalg.cholesky(spd(n, rng))
    B = rng.standard_normal((m, n))
    Bg = gla.colmajor(B)
    gla.rdiv_(Bg, gla.adjoint(gla.LowerTriangular(gla.colmajor(L))))
    ref = O.trsm_rlt(B, L)
    assert rel(gla.to_numpy(Bg), ref) <= 10 * n * eps_of(np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,block", [(3000, 768), (5000, 512)])
def test_side_cu_reserve_is_bitwise_neutral(rng, dtype, n, block):
    """The look-ahead driver with its trailing SYRK on a CU-masked stream and
    the chain on the handle's own stream (the default) computes the same
    bits as with both on unmasked streams (side_cu_reserve 0), also when
    the caller's stream is not the default one."""
    gla = _gla()
    A = spd(n, rng).astype(dtype)
    outs = []
    for reserve in (gla.get_tuning("side_cu_reserve"), 0):
        with gla.tuning(chol_block=block, side_cu_reserve=reserve):
            X = gla.colmajor(A)
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                gla.cholRecursive_(X, gla.Val.L)
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            outs.append(gla.to_numpy(X))
    assert np.array_equal(outs[0], outs[1])
    assert rel(np.tril(outs[0]), np.linalg.cholesky(A.astype(np.float64))) <= 10 * n * eps_of(dtype)


def test_tuning_defaults_and_restore():
    gla = _gla()
    assert gla.get_tuning("chol_block") == -1
    assert gla.get_tuning("gemm_dl") == 2
    assert gla.get_tuning("chol_psolve") == 1
    assert gla.get_tuning("side_cu_reserve") == torch.cuda.get_device_properties(0).multi_processor_count // 8
    with gla.tuning(chol_block=256, side_cu_reserve=0):
        assert gla.get_tuning("chol_block") == 256
        assert gla.get_tuning("side_cu_reserve") == 0
    assert gla.get_tuning("chol_block") == -1
    with pytest.raises(gla.ArgumentError):
        gla.get_tuning("no_such_knob")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,block", [(1100, 256), (2000, 384), (2500, 768), (3000, 640),
                                     (3000, 1024)])
def test_fused_panel_solve_matches_recursive(rng, dtype, n, block):
    """The look-ahead driver's panels solved in one launch each
    (chol_psolve 1, the default; blocks up to 768 wide, wider ones take the
    recursive solve either way) against the recursive rdiv! (0) and
    LAPACK; ragged last steps and row counts that are not multiples of the
    64-row strips included; the strict upper triangle untouched."""
    gla = _gla()
    A = spd(n, rng).astype(dtype)
    outs = []
    for ps in (1, 0):
        with gla.tuning(chol_block=block, chol_psolve=ps):
            out, _ = _run(lambda X: gla.cholRecursive_(X, gla.Val.L), A, dtype)
        outs.append(out)
    tol = 10 * n * eps_of(dtype)
    assert rel(np.tril(outs[0]), np.tril(outs[1])) <= tol
    assert rel(np.tril(outs[0]), np.linalg.cholesky(A.astype(np.float64))) <= tol
    assert np.array_equal(np.triu(outs[0], 1), np.triu(A, 1))
