"""ctypes front end of the CPU oracle (oracle/gla_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the benchmark's cpu_baseline leg, always as the checker or the CPU
baseline, never as the product path.

Every function takes and returns numpy arrays in column-major (Fortran)
order and copies its inputs, so callers keep their originals.  The suffix
of the C symbol is picked from the dtype (float32 -> _s, float64 -> _d);
``*_q`` helpers compute in __float128 (the stand-in for the reference's
BigFloat runs) from float64 inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libgla_oracle.so")
_lib = None

_i64 = C.c_int64
_p = C.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        for sfx, ft in (("d", C.c_double), ("s", C.c_float)):
            g = lambda n: getattr(L, f"{n}_{sfx}")  # noqa: E731
            g("orc_chol_unblocked").argtypes = [_i64, _p, _i64]
            g("orc_chol_unblocked").restype = _i64
            g("orc_chol_blocked").argtypes = [_i64, _p, _i64, _i64]
            g("orc_chol_blocked").restype = _i64
            g("orc_chol_recursive").argtypes = [_i64, _p, _i64, _i64]
            g("orc_chol_recursive").restype = _i64
            g("orc_syr_l").argtypes = [_i64, ft, _p, _i64, _p, _i64]
            g("orc_syrk_l").argtypes = [_i64, _i64, ft, _p, _i64, _p, _i64]
            g("orc_trsm_rlt").argtypes = [_i64, _i64, _p, _i64, _p, _i64]
            g("orc_reflector").argtypes = [_i64, _p]
            g("orc_reflector").restype = ft
            g("orc_reflector_apply").argtypes = [_i64, _i64, _p, ft, _p, _i64]
            g("orc_qr_unblocked").argtypes = [_i64, _i64, _p, _i64, _p]
            g("orc_qr_tblock").argtypes = [_i64, _i64, _p, _i64, _p, _p, _i64]
            g("orc_lmul_unit_lower").argtypes = [_i64, _i64, _p, _i64, _p, _i64, ft]
            g("orc_wy_apply").argtypes = [C.c_int, _i64, _i64, _p, _i64, _i64, _p, _i64,
                                          _p, _i64, _p, _i64]
            g("orc_qr_blocked").argtypes = [_i64, _i64, _p, _i64, _i64, _p, _p, _i64]
            g("orc_ldlt_lower").argtypes = [_i64, _p, _i64]
            g("orc_ldlt_upper").argtypes = [_i64, _p, _i64]
            g("orc_ldlt_lower_blocked").argtypes = [_i64, _p, _i64, _i64]
            g("orc_ldlt_upper_blocked").argtypes = [_i64, _p, _i64, _i64]
            g("orc_ldlt_solve").argtypes = [C.c_int, _i64, _i64, _p, _i64, _p, _i64]
            g("orc_symtri_lower").argtypes = [_i64, _p, _i64, _p]
            g("orc_symtri_upper").argtypes = [_i64, _p, _i64, _p]
            g("orc_bidiag").argtypes = [_i64, _i64, _p, _i64, _p, _p]
        L.orq_chol_recursive.argtypes = [_i64, _p, _i64, _i64]
        L.orq_chol_recursive.restype = _i64
        L.orq_chol_unblocked.argtypes = [_i64, _p, _i64]
        L.orq_chol_unblocked.restype = _i64
        L.orq_qr_unblocked.argtypes = [_i64, _i64, _p, _i64, _p]
        for sfx, rt in (("z", C.c_double), ("c", C.c_float)):
            g = lambda n: getattr(L, f"{n}_{sfx}")  # noqa: E731
            for nm in ("orz_chol_unblocked",):
                g(nm).argtypes = [_i64, _p, _i64]
                g(nm).restype = _i64
            g("orz_chol_blocked").argtypes = [_i64, _p, _i64, _i64]
            g("orz_chol_blocked").restype = _i64
            g("orz_chol_recursive").argtypes = [_i64, _p, _i64, _i64]
            g("orz_chol_recursive").restype = _i64
            g("orz_herk_l").argtypes = [_i64, _i64, rt, _p, _i64, _p, _i64]
            g("orz_her_l").argtypes = [_i64, rt, _p, _p, _i64]
            g("orz_trsm_rlc").argtypes = [_i64, _i64, _p, _i64, _p, _i64]
            g("orz_qr_unblocked").argtypes = [_i64, _i64, _p, _i64, _p]
        L.orq_symtri.argtypes = [C.c_int, _i64, _p, _i64, _p]
        L.orq_bidiag.argtypes = [_i64, _i64, _p, _i64, _p, _p]
        L.orq_chol_residual.argtypes = [_i64, _p, _i64, _p, _i64]
        L.orq_chol_residual.restype = C.c_double
        _lib = L
    return _lib


def _sfx(a: np.ndarray) -> str:
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "s"
    if a.dtype == np.complex128:
        return "z"
    if a.dtype == np.complex64:
        return "c"
    raise TypeError(f"oracle supports float32/float64, got {a.dtype}")


def _f(a, dtype=None) -> np.ndarray:
    a = np.array(a, dtype=dtype, order="F", copy=True)
    if a.ndim == 1:
        a = a.reshape(-1)
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p)


def _fn(name: str, a: np.ndarray):
    return getattr(lib(), f"{name}_{_sfx(a)}")


# --------------------------------------------------------------- Cholesky --
def chol_unblocked(A):
    """cholUnblocked!(copy(A), Val{:L}) -> (A_out, info); src/cholesky.jl:3-15."""
    A = _f(A)
    info = _fn("orc_chol_unblocked", A)(A.shape[0], _ptr(A), A.shape[0])
    return A, int(info)


def chol_blocked(A, blocksize: int):
    """cholBlocked!(copy(A), Val{:L}, blocksize) -> (A_out, info); src/cholesky.jl:17-35."""
    A = _f(A)
    info = _fn("orc_chol_blocked", A)(A.shape[0], _ptr(A), A.shape[0], int(blocksize))
    return A, int(info)


def chol_recursive(A, cutoff: int = 1):
    """cholRecursive!(copy(A), Val{:L}, cutoff) -> (A_out, info); src/cholesky.jl:37-55."""
    A = _f(A)
    info = _fn("orc_chol_recursive", A)(A.shape[0], _ptr(A), A.shape[0], int(cutoff))
    return A, int(info)


def chol_recursive_q(A, cutoff: int = 1):
    """cholRecursive! evaluated in __float128 from float64 input, rounded back."""
    A = _f(A, np.float64)
    info = lib().orq_chol_recursive(A.shape[0], _ptr(A), A.shape[0], int(cutoff))
    return A, int(info)


def chol_unblocked_q(A):
    A = _f(A, np.float64)
    info = lib().orq_chol_unblocked(A.shape[0], _ptr(A), A.shape[0])
    return A, int(info)


def chol_residual_q(A, L) -> float:
    """‖A − LLᵀ‖_F/‖A‖_F in __float128 (lower triangles of A and L)."""
    A = _f(A, np.float64)
    L = _f(L, np.float64)
    return float(lib().orq_chol_residual(A.shape[0], _ptr(A), A.shape[0], _ptr(L), L.shape[0]))


def syr_l(A, a, alpha):
    """rankUpdate!(Hermitian(A,:L), a, α) generic rank-1; src/juliaBLAS.jl:53-63."""
    A = _f(A)
    a = _f(a, A.dtype)
    _fn("orc_syr_l", A)(A.shape[0], alpha, _ptr(a), 1, _ptr(A), A.shape[0])
    return A


def syrk_l(C_, A, alpha):
    """rankUpdate!(Hermitian(C,:L), A, α) generic rank-k; src/juliaBLAS.jl:89-112."""
    C_ = _f(C_)
    A = _f(A, C_.dtype)
    if A.ndim == 1:
        A = A.reshape(-1, 1, order="F")
    _fn("orc_syrk_l", C_)(C_.shape[0], A.shape[1], alpha, _ptr(A), A.shape[0], _ptr(C_),
                          C_.shape[0])
    return C_


def trsm_rlt(B, L):
    """rdiv!(B, LowerTriangular(L)') -> B L^{-T}."""
    B = _f(B)
    L = _f(L, B.dtype)
    _fn("orc_trsm_rlt", B)(B.shape[0], B.shape[1], _ptr(L), L.shape[0], _ptr(B), B.shape[0])
    return B


# --------------------------------------------------------------------- QR --
def reflector(x):
    """LinearAlgebra.reflector!(copy(x)) -> (x_out, τ)."""
    x = _f(x)
    tau = _fn("orc_reflector", x)(x.shape[0], _ptr(x))
    return x, float(tau)


def reflector_apply(x, tau, A):
    """LinearAlgebra.reflectorApply!(x, τ, copy(A)) (left)."""
    A = _f(A)
    x = _f(x, A.dtype)
    if x.shape[0] != A.shape[0]:
        raise ValueError("reflector must have same length as first dimension of matrix")
    _fn("orc_reflector_apply", A)(A.shape[0], A.shape[1], _ptr(x), tau, _ptr(A), A.shape[0])
    return A


def qr_unblocked(A):
    """qrUnblocked!(copy(A)) -> (factors, τ); src/qr.jl:86-111."""
    A = _f(A)
    m, n = A.shape
    tau = np.zeros(min(m, n), dtype=A.dtype)
    _fn("orc_qr_unblocked", A)(m, n, _ptr(A), m, _ptr(tau))
    return A, tau


def qr_unblocked_q(A):
    A = _f(A, np.float64)
    m, n = A.shape
    tau = np.zeros(min(m, n), dtype=np.float64)
    lib().orq_qr_unblocked(m, n, _ptr(A), m, _ptr(tau))
    return A, tau


def qr_blocked(A, blocksize: int = 12, return_work: bool = False):
    """qrBlocked!(copy(A), blocksize) -> (factors, τ), or (factors, τ, work)
    with the blocksize × n workspace as the call leaves it (zero-initialised);
    src/qr.jl:113-146."""
    A = _f(A)
    m, n = A.shape
    tau = np.zeros(min(m, n), dtype=A.dtype)
    work = np.zeros((blocksize, max(n, 1)), dtype=A.dtype, order="F")
    _fn("orc_qr_blocked", A)(m, n, _ptr(A), m, int(blocksize), _ptr(tau), _ptr(work),
                             blocksize)
    return (A, tau, work) if return_work else (A, tau)


def qr_tblock(F, tau):
    """F[Tuple{:QBlocked}].T of a QR2(F, τ); src/qr.jl:64-83."""
    F = _f(F)
    tau = _f(tau, F.dtype)
    m, n = F.shape
    k = min(m, n)
    T = np.zeros((k, k), dtype=F.dtype, order="F")
    _fn("orc_qr_tblock", F)(m, n, _ptr(F), m, _ptr(tau), _ptr(T), max(k, 1))
    return T


def wy_apply(V, T, A, adjoint: bool):
    """lmul!(H, copy(A), M) or lmul!(H', copy(A), M) for H = HouseholderBlock(V, T);
    src/householder.jl:82-115,119-157."""
    A = _f(A)
    V = _f(V, A.dtype)
    T = _f(T, A.dtype)
    mA, nA = A.shape
    if V.shape[0] != mA:
        raise ValueError("DimensionMismatch")
    b = min(mA, V.shape[1])
    M = np.zeros((max(b, 1), max(nA, 1)), dtype=A.dtype, order="F")
    _fn("orc_wy_apply", A)(1 if adjoint else 0, mA, nA, _ptr(V), mA, V.shape[1], _ptr(T),
                           T.shape[0], _ptr(A), mA, _ptr(M), M.shape[0])
    return A


def lmul_unit_lower(L, B, alpha):
    """lmul!(UnitLowerTriangular(L), copy(B), α); src/juliaBLAS.jl:154-166."""
    B = _f(B)
    L = _f(L, B.dtype)
    _fn("orc_lmul_unit_lower", B)(B.shape[0], B.shape[1], _ptr(L), L.shape[0], _ptr(B),
                                  B.shape[0], alpha)
    return B


# ------------------------------------------------------------ complex --
def zchol(A, variant: str = "recursive", param: int = 1):
    """cholUnblocked!/cholBlocked!/cholRecursive! of a complex matrix
    (src/cholesky.jl:3-55, with her!/herk! for the rank updates) ->
    (A_out, first index of a negative real pivot or 0)."""
    A = _f(A)
    n = A.shape[0]
    if variant == "unblocked":
        info = _fn("orz_chol_unblocked", A)(n, _ptr(A), max(n, 1))
    elif variant == "blocked":
        info = _fn("orz_chol_blocked", A)(n, _ptr(A), max(n, 1), int(param))
    else:
        info = _fn("orz_chol_recursive", A)(n, _ptr(A), max(n, 1), int(param))
    return A, int(info)


def zherk_l(C_, A, alpha):
    """rankUpdate!(Hermitian(C, :L), A, α) for complex (herk!)."""
    C_ = _f(C_)
    A = _f(A, C_.dtype)
    if A.ndim == 1:
        A = A.reshape(-1, 1, order="F")
    _fn("orz_herk_l", C_)(C_.shape[0], A.shape[1], alpha, _ptr(A), A.shape[0], _ptr(C_),
                          C_.shape[0])
    return C_


def zqr_unblocked(A):
    """qrUnblocked!(copy(A)) for complex A -> (factors, τ); src/qr.jl:86-111."""
    A = _f(A)
    m, n = A.shape
    tau = np.zeros(min(m, n), dtype=A.dtype)
    _fn("orz_qr_unblocked", A)(m, n, _ptr(A), m, _ptr(tau))
    return A, tau


# ------------------------------------------------------------------- LDLᵀ --
def ldlt_default_blocksize(dtype) -> int:
    """ldlt!'s default blocksize max(1, 128 ÷ sizeof(T)) (src/ldlt.jl:150)."""
    return max(1, 128 // np.dtype(dtype).itemsize)


def ldlt(A, uplo: str = "L", blocksize: int | None = None):
    """ldlt!(Hermitian(copy(A), uplo), blocksize).data (src/ldlt.jl:150-157):
    the blocked factorisation :80-148, only the `uplo` triangle read and
    written.  blocksize ≥ n runs the unblocked form (:16-62)."""
    A = _f(A)
    n = A.shape[0]
    bs = ldlt_default_blocksize(A.dtype) if blocksize is None else int(blocksize)
    name = "orc_ldlt_lower_blocked" if uplo == "L" else "orc_ldlt_upper_blocked"
    _fn(name, A)(n, _ptr(A), max(n, 1), bs)
    return A


def ldlt_unblocked(A, uplo: str = "L"):
    """_ldlt_lower!(copy(A)) / _ldlt_upper!(copy(A)); src/ldlt.jl:16-62."""
    A = _f(A)
    n = A.shape[0]
    _fn("orc_ldlt_lower" if uplo == "L" else "orc_ldlt_upper", A)(n, _ptr(A), max(n, 1))
    return A


def ldlt_solve(F, B, uplo: str = "L"):
    """ldiv!(F::LDLt, copy(B)) for the factor data F; src/ldlt.jl:225-228."""
    B = _f(B)
    F = _f(F, B.dtype)
    vec = B.ndim == 1
    if vec:
        B = B.reshape(-1, 1, order="F")
    n = F.shape[0]
    _fn("orc_ldlt_solve", B)(1 if uplo == "U" else 0, n, B.shape[1], _ptr(F), max(n, 1), _ptr(B),
                             max(n, 1))
    return B.reshape(-1) if vec else B


def ldlt_factors(F, uplo: str = "L"):
    """(L, d) of factor data F: L unit lower (L = U' for uplo 'U'), d the
    diagonal of D (the getproperty of src/ldlt.jl:208-222)."""
    F = np.asarray(F)
    d = np.diag(F).copy()
    L = np.tril(F, -1) if uplo == "L" else np.triu(F, 1).T
    L = L + np.eye(F.shape[0], dtype=F.dtype)
    return L, d


# ---------------------------------------------------- two-sided reductions --
def symtri(A, uplo: str = "L"):
    """symtri!(Hermitian(copy(A), uplo)) -> (AS, τ) for real A: symtriLower!
    / symtriUpper! (src/eigenSelfAdjoint.jl:450-564).  AS holds the
    tridiagonal (diagonal and sub/superdiagonal) and the reflectors; τ has
    n − 1 entries (the last 0)."""
    A = _f(A)
    n = A.shape[0]
    tau = np.zeros(max(n - 1, 0), dtype=A.dtype)
    name = "orc_symtri_lower" if uplo == "L" else "orc_symtri_upper"
    _fn(name, A)(n, _ptr(A), max(n, 1), _ptr(tau) if n > 1 else None)
    return A, tau


def symtri_quad(A, uplo: str = "L"):
    """symtri! evaluated in __float128 from float64 input, rounded back."""
    A = _f(A, np.float64)
    n = A.shape[0]
    tau = np.zeros(max(n - 1, 1), dtype=np.float64)
    lib().orq_symtri(1 if uplo == "U" else 0, n, _ptr(A), max(n, 1), _ptr(tau))
    return A, tau[:max(n - 1, 0)]


def bidiag_quad(A):
    """bidiagonalize! evaluated in __float128 from float64 input, rounded back."""
    A = _f(A, np.float64)
    m, n = A.shape
    k = min(m, n)
    nl, nr = (k, max(k - 1, 0)) if m >= n else (max(k - 1, 0), k)
    taul = np.zeros(max(nl, 1))
    taur = np.zeros(max(nr, 1))
    lib().orq_bidiag(m, n, _ptr(A), max(m, 1), _ptr(taul), _ptr(taur))
    return A, taul[:nl], taur[:nr]


def symtri_q(AS, tau, uplo: str = "L"):
    """The orthogonal Q of symtri!'s reflectors (float64 numpy): Q = H₁H₂⋯
    with Qᵀ A Q tridiagonal.  uplo 'L': H_k reflects rows k+1: with
    v = [1, AS[k+2:, k]]; 'U': H_k reflects rows 1:n−k with v read upwards
    from AS[n−k−1, n−k] (= 1) through AS[0, n−k]."""
    AS = np.asarray(AS, dtype=np.float64)
    n = AS.shape[0]
    Q = np.eye(n)
    for k in range(n - 2):
        v = np.zeros(n)
        if uplo == "L":
            v[k + 1] = 1.0
            v[k + 2:] = AS[k + 2:, k]
        else:
            m = n - k - 1
            v[m - 1] = 1.0
            v[:m - 1] = AS[:m - 1, m]
        Q = Q - float(tau[k]) * np.outer(Q @ v, v)
    return Q


def symtri_diagonals(AS, uplo: str = "L"):
    """(d, e): the tridiagonal's diagonal and off-diagonal."""
    AS = np.asarray(AS)
    return np.diag(AS).copy(), (np.diag(AS, -1) if uplo == "L" else np.diag(AS, 1)).copy()


def bidiag(A):
    """bidiagonalize!(copy(A)) -> (A_out, τl, τr) for real A (src/svd.jl:328-381)."""
    A = _f(A)
    m, n = A.shape
    k = min(m, n)
    nl, nr = (k, max(k - 1, 0)) if m >= n else (max(k - 1, 0), k)
    taul = np.zeros(max(nl, 1), dtype=A.dtype)
    taur = np.zeros(max(nr, 1), dtype=A.dtype)
    _fn("orc_bidiag", A)(m, n, _ptr(A), max(m, 1), _ptr(taul), _ptr(taur))
    return A, taul[:nl], taur[:nr]


def bidiag_q(F, taul, taur):
    """(Ql, Qr, B) of bidiagonalize!'s output (float64 numpy): Ql (m×m) and
    Qr (n×n) the products of the left and right reflectors, B = Qlᵀ A Qr
    the bidiagonal (upper for m ≥ n, lower otherwise)."""
    F = np.asarray(F, dtype=np.float64)
    m, n = F.shape
    Ql, Qr = np.eye(m), np.eye(n)
    if m >= n:
        for i, t in enumerate(taul):
            v = np.zeros(m)
            v[i] = 1.0
            v[i + 1:] = F[i + 1:, i]
            Ql = Ql - float(t) * np.outer(Ql @ v, v)
        for i, t in enumerate(taur):
            v = np.zeros(n)
            v[i + 1] = 1.0
            v[i + 2:] = F[i, i + 2:]
            Qr = Qr - float(t) * np.outer(Qr @ v, v)
        B = np.diag(np.diag(F)) + np.diag(np.diag(F, 1), 1)
        B = np.vstack([B, np.zeros((m - n, n))])
    else:
        for i, t in enumerate(taur):
            v = np.zeros(n)
            v[i] = 1.0
            v[i + 1:] = F[i, i + 1:]
            Qr = Qr - float(t) * np.outer(Qr @ v, v)
        for i, t in enumerate(taul):
            v = np.zeros(m)
            v[i + 1] = 1.0
            v[i + 2:] = F[i + 2:, i]
            Ql = Ql - float(t) * np.outer(Ql @ v, v)
        B = np.diag(np.diag(F)) + np.diag(np.diag(F, -1), -1)
        B = np.hstack([B, np.zeros((m, n - m))])
    return Ql, Qr, B


# --------------------------------------------------------------- adjoints --
# No reference code: restated from the published formulas, in float64 numpy
# (general solves, no triangular shortcuts), for small sizes.
def thin_q(F, tau):
    """Q[:, :n] of the reflectors stored in F (m×n, m ≥ n) and τ, formed as
    H₁H₂⋯Hₙ [I; 0] with Hₖ = I − τₖ vₖvₖᵀ, vₖ = [0…0, 1, F[k+1:, k]]."""
    F = np.asarray(F, dtype=np.float64)
    m, n = F.shape
    E = np.zeros((m, n))
    E[:n, :n] = np.eye(n)
    for k in range(n - 1, -1, -1):
        v = np.zeros(m)
        v[k] = 1.0
        v[k + 1:] = F[k + 1:, k]
        E -= float(tau[k]) * np.outer(v, v @ E)
    return E


def chol_adjoint(L, Lbar, conv: str = "S"):
    """Ā of A = LLᵀ for the cotangent L̄ (Murray 2016, arXiv:1602.07527,
    eq. 10): P = Φ(Lᵀ tril(L̄)), Φ = lower triangle with halved diagonal;
    Ā = ½ L⁻ᵀ (P + Pᵀ) L⁻¹.  conv "S": that symmetric matrix; "L": the
    derivative with respect to the lower triangle (2Ā below the diagonal,
    Ā on it, 0 above)."""
    L = np.tril(np.asarray(L, dtype=np.float64))
    G = np.tril(np.asarray(Lbar, dtype=np.float64))
    P = np.tril(L.T @ G)
    P[np.diag_indices_from(P)] *= 0.5
    Y = np.linalg.solve(L.T, 0.5 * (P + P.T))     # L⁻ᵀ Psym
    Ab = np.linalg.solve(L.T, Y.T).T              # (L⁻ᵀ Yᵀ)ᵀ = Y L⁻¹
    if conv == "L":
        Ab = np.tril(2.0 * Ab) - np.diag(np.diag(Ab))
    return Ab


def qr_adjoint(F, tau, Rbar):
    """Ā of A = QR (m ≥ n, Q̄ = 0) for the cotangent R̄ (upper part read):
    M = R R̄ᵀ, Ā = Q copyltu(M) R⁻ᵀ with copyltu(M) = tril(M) + tril(M,−1)ᵀ
    (Walter & Lehmann 2018; Liao et al. 2019)."""
    F = np.asarray(F, dtype=np.float64)
    m, n = F.shape
    R = np.triu(F[:n, :n])
    M = R @ np.triu(np.asarray(Rbar, dtype=np.float64)).T
    Cm = np.tril(M) + np.tril(M, -1).T
    X = np.linalg.solve(R, Cm.T).T                # X Rᵀ = C
    return thin_q(F, tau) @ X
