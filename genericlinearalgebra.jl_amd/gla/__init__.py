"""Host-side mirror of GenericLinearAlgebra.jl's blocked factorisation API
over libgla_hip.so (gfx950).

Julia names map to Python as ``f!`` -> ``f_`` (in place); ``Val{:L}`` ->
``Val.L``; ``F[Tuple{:R}]`` -> ``F[R]``; ``F[Tuple{:QBlocked}]`` ->
``F[QBlocked]``; ``H'`` -> ``adjoint(H)``.  Matrices are device tensors in
column-major layout (``stride(0) == 1``), the storage the C ABI takes; use
:func:`colmajor` to make one.  Errors follow the reference: ``DomainError``
for a negative pivot (sqrt of a negative number), ``DimensionMismatch`` and
``ArgumentError`` as the reference throws them.

PyTorch supplies device memory and the current stream only; every
computation is a call into libgla_hip.so.
"""
from __future__ import annotations

import contextlib
import ctypes as C

import numpy as np
import torch

from ._lib import GLA_ERR_TIMEOUT, GLA_OK, build, exported_symbols, lib  # noqa: F401

__all__ = [
    "Val", "DomainError", "DimensionMismatch", "ArgumentError", "DeviceTimeout", "colmajor",
    "to_numpy",
    "LowerTriangular", "UpperTriangular", "UnitLowerTriangular", "Hermitian",
    "cholUnblocked_", "cholBlocked_", "cholRecursive_", "rankUpdate_", "rdiv_", "lmul_",
    "QR2", "R", "QBlocked", "HouseholderBlock", "Adjoint", "adjoint", "qrUnblocked_",
    "qrBlocked_", "gemm_", "LDLt", "ldlt_", "ldlt", "SymmetricTridiagonalFactorization",
    "symtri_", "BidiagonalFactorization", "bidiagonalize_",
]


class DomainError(ValueError):
    """sqrt of a negative pivot (Julia's DomainError)."""


class DimensionMismatch(ValueError):
    pass


class DeviceTimeout(RuntimeError):
    """A bounded device-side wait gave up (GLA_ERR_TIMEOUT); the call's
    outputs are not valid."""


class ArgumentError(ValueError):
    pass


class Val:
    """``Val{:L}`` / ``Val{:U}``."""
    L = "L"
    U = "U"


class _Tag:
    def __init__(self, name):
        self.name = name

    def __repr__(self):
        return f"Tuple{{:{self.name}}}"


R = _Tag("R")
QBlocked = _Tag("QBlocked")


# ------------------------------------------------------------------ plumbing
_handles: dict[int, int] = {}


def _handle(dev: int) -> int:
    h = _handles.get(dev)
    if h is None:
        with torch.cuda.device(dev):
            hp = C.c_void_p()
            _check(lib().gla_handle_create(C.byref(hp)), "gla_handle_create")
            h = hp.value
        _handles[dev] = h
    return h


def set_tuning(key: str, value: int, device: int = 0):
    """Sets a tuning knob of the device's handle (see gla_set_tuning)."""
    _check(lib().gla_set_tuning(_handle(device), key.encode(), int(value)), "gla_set_tuning")


def get_tuning(key: str, device: int = 0) -> int:
    """Reads a tuning knob of the device's handle (gla_get_tuning)."""
    v = C.c_int64(0)
    _check(lib().gla_get_tuning(_handle(device), key.encode(), C.byref(v)), "gla_get_tuning")
    return int(v.value)


@contextlib.contextmanager
def tuning(device: int = 0, **knobs):
    """Sets knobs for the duration of a with-block and restores the values
    they had before it, also when the block raises."""
    old = {k: get_tuning(k, device) for k in knobs}
    try:
        for k, v in knobs.items():
            set_tuning(k, v, device)
        yield
    finally:
        for k, v in old.items():
            set_tuning(k, v, device)


def _check(rc: int, what: str):
    if rc != GLA_OK:
        msg = lib().gla_last_error_string().decode()
        if rc < 0 and rc > -100:
            raise ArgumentError(f"{what}: invalid argument {-rc}: {msg}")
        if rc == GLA_ERR_TIMEOUT:
            raise DeviceTimeout(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def _sfx(A: torch.Tensor) -> str:
    if A.dtype == torch.float64:
        return "d"
    if A.dtype == torch.float32:
        return "s"
    if A.dtype == torch.complex128:
        return "z"
    if A.dtype == torch.complex64:
        return "c"
    raise TypeError(f"element type {A.dtype} not supported (Float32/Float64/ComplexF32/ComplexF64)")


def _mat(A: torch.Tensor, name="A") -> int:
    """Validates a column-major device matrix; returns its leading dimension."""
    if not isinstance(A, torch.Tensor) or A.dim() != 2:
        raise TypeError(f"{name} must be a 2-D torch tensor")
    if not A.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    m, n = A.shape
    if n > 1 and A.stride(0) != 1:
        raise TypeError(f"{name} must be column-major (stride(0) == 1); see colmajor()")
    ld = A.stride(1) if n > 1 else max(1, m)
    if m > 1 and A.stride(0) != 1:
        raise TypeError(f"{name} must be column-major")
    return max(ld, 1)


def _vec(x: torch.Tensor, n: int, dtype, name="τ") -> torch.Tensor:
    if x.dim() != 1 or x.numel() < n or x.dtype != dtype or not x.is_cuda or x.stride(0) != 1:
        raise TypeError(f"{name} must be a contiguous device vector of length >= {n}")
    return x


def _stream(A: torch.Tensor):
    return C.c_void_p(torch.cuda.current_stream(A.device).cuda_stream)


def colmajor(x, dtype=None, device="cuda") -> torch.Tensor:
    """Column-major device copy of a numpy array or tensor (Julia layout)."""
    if isinstance(x, torch.Tensor):
        t = x.detach()
    else:
        t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    if t.dim() == 1:
        return t.to(device).contiguous()
    m, n = t.shape
    out = torch.empty((n, m), dtype=t.dtype, device=device).t()
    out.copy_(t)
    return out


def empty_colmajor(m: int, n: int, dtype=torch.float64, device="cuda") -> torch.Tensor:
    return torch.empty((n, m), dtype=dtype, device=device).t()


def to_numpy(A: torch.Tensor) -> np.ndarray:
    return np.asfortranarray(A.detach().cpu().numpy())


# ------------------------------------------------------------- wrappers ----
class _Wrap:
    def __init__(self, data):
        self.data = data

    def __repr__(self):
        return f"{type(self).__name__}({self.data!r})"


class LowerTriangular(_Wrap):
    pass


class UpperTriangular(_Wrap):
    pass


class UnitLowerTriangular(_Wrap):
    pass


class Hermitian:
    """``Hermitian(C, :L)``: only the ``uplo`` triangle of ``data`` is used."""

    def __init__(self, data, uplo="U"):
        self.data = data
        self.uplo = uplo.strip(":").upper()


class Adjoint:
    def __init__(self, parent):
        self.parent = parent


def adjoint(x):
    if isinstance(x, Adjoint):
        return x.parent
    return Adjoint(x)


# ------------------------------------------------------------- Cholesky ----
def _uplo(v) -> str:
    u = str(v).strip(":").upper()
    if u not in ("L", "U"):
        raise ArgumentError("uplo must be Val.L or Val.U")
    return u


_last_info: dict[int, int] = {}


def last_pivot_info(device: int = 0) -> int:
    """The info of the last Cholesky call on the device: the 1-based index
    of the first pivot whose real part was negative, or 0 (for complex
    matrices the reference raises no error there)."""
    return _last_info.get(device, 0)


def _chol(A, uplo, fn, *extra):
    ld = _mat(A)
    m, n = A.shape
    if m != n:
        raise DimensionMismatch(f"matrix is not square: dimensions are {tuple(A.shape)}")
    info = C.c_int64(0)
    f = getattr(lib(), f"{fn}_{_sfx(A)}")
    with torch.cuda.device(A.device):
        rc = f(_handle(A.device.index), _uplo(uplo).encode(), n, C.c_void_p(A.data_ptr()), ld,
               *extra, C.byref(info), _stream(A))
    _check(rc, fn)
    if info.value > 0 and not A.is_complex():
        # complex pivots take the complex square root (no error), as in
        # the reference; info is kept on the tensor for inspection
        raise DomainError(f"sqrt of a negative pivot at position {info.value}")
    _last_info[A.device.index] = info.value
    return A


def cholUnblocked_(A, uplo=Val.L):
    """cholUnblocked!(A, Val{:L}) — src/cholesky.jl:3-15.  Returns A."""
    return _chol(A, uplo, "gla_chol_unblocked")


def cholBlocked_(A, uplo, blocksize: int):
    """cholBlocked!(A, Val{:L}, blocksize) — src/cholesky.jl:17-35.  Returns A."""
    if int(blocksize) < 1:
        raise ArgumentError("blocksize must be positive")
    return _chol(A, uplo, "gla_chol_blocked", int(blocksize))


def cholRecursive_(A, uplo=Val.L, cutoff: int = 1):
    """cholRecursive!(A, Val{:L}, cutoff) — src/cholesky.jl:37-55.
    Returns LowerTriangular(A) (UpperTriangular for Val.U)."""
    _chol(A, uplo, "gla_chol_recursive", int(cutoff))
    return LowerTriangular(A) if _uplo(uplo) == "L" else UpperTriangular(A)


def rankUpdate_(Cm: Hermitian, A, alpha):
    """rankUpdate!(C::Hermitian, A, α) — src/juliaBLAS.jl:89-112 (and the
    rank-1 method :53-63 when A is a vector).  Returns C."""
    Cd = Cm.data
    ldc = _mat(Cd, "C")
    n = Cd.shape[0]
    if Cd.shape[1] != n:
        raise DimensionMismatch("C must be square")
    if A.dim() == 1:
        A = A.reshape(-1, 1)
    if A.shape[0] != n:
        raise DimensionMismatch("first dimension of A has wrong size")
    lda = _mat(A, "A")
    f = getattr(lib(), f"gla_rank_update_{_sfx(Cd)}")
    if Cd.is_complex() and complex(alpha).imag != 0:
        raise ArgumentError("α must be real for a Hermitian rank update")
    with torch.cuda.device(Cd.device):
        rc = f(_handle(Cd.device.index), Cm.uplo.encode(), n, A.shape[1], float(complex(alpha).real),
               C.c_void_p(A.data_ptr()), lda, C.c_void_p(Cd.data_ptr()), ldc, _stream(Cd))
    _check(rc, "gla_rank_update")
    return Cm


def rdiv_(B, Lt):
    """rdiv!(B, LowerTriangular(L)') as called at src/cholesky.jl:28,48:
    B <- B L⁻ᵀ.  ``Lt`` is ``adjoint(LowerTriangular(L))``."""
    if not (isinstance(Lt, Adjoint) and isinstance(Lt.parent, LowerTriangular)):
        raise TypeError("rdiv_ supports B / LowerTriangular(L)' only")
    L = Lt.parent.data
    ldb, ldl = _mat(B, "B"), _mat(L, "L")
    m, n = B.shape
    if L.shape != (n, n):
        raise DimensionMismatch("L must be n×n with n = size(B, 2)")
    f = getattr(lib(), f"gla_trsm_rlt_{_sfx(B)}")
    with torch.cuda.device(B.device):
        rc = f(_handle(B.device.index), m, n, C.c_void_p(L.data_ptr()), ldl,
               C.c_void_p(B.data_ptr()), ldb, _stream(B))
    _check(rc, "gla_trsm_rlt")
    return B


def gemm_(transa: str, transb: str, alpha, A, B, beta, Cm):
    """C = α op(A) op(B) + β C on the MFMA engine (mul!(C, A, B, α, β))."""
    lda, ldb, ldc = _mat(A, "A"), _mat(B, "B"), _mat(Cm, "C")
    m, n = Cm.shape
    k = A.shape[1] if transa == "N" else A.shape[0]
    f = getattr(lib(), f"gla_gemm_{_sfx(Cm)}")
    with torch.cuda.device(Cm.device):
        rc = f(_handle(Cm.device.index), transa.encode(), transb.encode(), m, n, k, float(alpha),
               C.c_void_p(A.data_ptr()), lda, C.c_void_p(B.data_ptr()), ldb, float(beta),
               C.c_void_p(Cm.data_ptr()), ldc, _stream(Cm))
    _check(rc, "gla_gemm")
    return Cm


# ------------------------------------------------------------------- QR ----
class HouseholderBlock:
    """Compact WY block I − V T Vᴴ (src/householder.jl:8-11).  ``T`` is an
    UpperTriangular wrapper of a k×k device matrix."""

    def __init__(self, V, T):
        self.V = V
        self.T = T if isinstance(T, UpperTriangular) else UpperTriangular(T)

    def adjoint(self):
        return Adjoint(self)

    def __matmul__(self, A):
        return _wy_mul(self, A, False)

    __mul__ = __matmul__


def _adj_matmul(self, A):
    if isinstance(self.parent, HouseholderBlock):
        return _wy_mul(self.parent, A, True)
    raise TypeError("unsupported adjoint product")


Adjoint.__matmul__ = _adj_matmul
Adjoint.__mul__ = _adj_matmul


def lmul_(H, A, M=None, alpha=None):
    """lmul!(H, A, M) / lmul!(H', A, M) for a HouseholderBlock
    (src/householder.jl:82-115,119-157), or lmul!(UnitLowerTriangular(L), B, α)
    (src/juliaBLAS.jl:154-166).  Returns A."""
    if isinstance(H, UnitLowerTriangular):
        if alpha is None:
            alpha = M if M is not None else 1.0
        L = H.data
        ldb, ldl = _mat(A, "B"), _mat(L, "L")
        m, n = A.shape
        if L.shape[0] < m or L.shape[1] < m:
            raise DimensionMismatch("L must be at least m×m")
        f = getattr(lib(), f"gla_lmul_unit_lower_{_sfx(A)}")
        with torch.cuda.device(A.device):
            rc = f(_handle(A.device.index), m, n, C.c_void_p(L.data_ptr()), ldl,
                   C.c_void_p(A.data_ptr()), ldb, float(alpha), _stream(A))
        _check(rc, "gla_lmul_unit_lower")
        return A
    trans = "N"
    if isinstance(H, Adjoint):
        H = H.parent
        trans = "T"
    if not isinstance(H, HouseholderBlock):
        raise TypeError("lmul_ expects a HouseholderBlock, its adjoint or UnitLowerTriangular")
    V, T = H.V, H.T.data
    ldv, ldt, lda = _mat(V, "V"), _mat(T, "T"), _mat(A, "A")
    mA, nA = A.shape
    if V.shape[0] != mA:
        raise DimensionMismatch("")
    b = min(mA, V.shape[1])
    if M is not None:
        ldm = _mat(M, "M")
        if M.shape[0] < b or M.shape[1] < nA:
            raise DimensionMismatch("workspace M is too small")
        mp = C.c_void_p(M.data_ptr())
    else:
        ldm, mp = max(b, 1), C.c_void_p(None)
    f = getattr(lib(), f"gla_wy_apply_{_sfx(A)}")
    with torch.cuda.device(A.device):
        if A.is_complex():        # Hᴴ for complex blocks; the workspace is internal
            rc = f(_handle(A.device.index), ("C" if trans == "T" else "N").encode(), mA, nA,
                   C.c_void_p(V.data_ptr()), ldv, V.shape[1], C.c_void_p(T.data_ptr()), ldt,
                   C.c_void_p(A.data_ptr()), lda, _stream(A))
        else:
            rc = f(_handle(A.device.index), trans.encode(), mA, nA, C.c_void_p(V.data_ptr()),
                   ldv, V.shape[1], C.c_void_p(T.data_ptr()), ldt, C.c_void_p(A.data_ptr()), lda,
                   mp, ldm, _stream(A))
    _check(rc, "gla_wy_apply")
    return A


def _wy_mul(H, A, adj):
    """``H * A`` / ``H' * A`` (src/householder.jl:116-117,158-159): copies A."""
    out = colmajor(A, device=A.device)
    return lmul_(Adjoint(H) if adj else H, out)


class QR2:
    """Result of qrUnblocked!/qrBlocked! (src/qr.jl:6-16): ``factors`` holds R
    in the upper triangle and the reflectors below it, ``τ`` their loadings."""

    def __init__(self, factors, tau):
        self.factors = factors
        self.τ = tau

    @property
    def tau(self):
        return self.τ

    def size(self, *i):
        s = tuple(self.factors.shape)
        return s if not i else tuple(s[j - 1] for j in i)

    def __getitem__(self, key):
        m, n = self.factors.shape
        if key is R:                                            # src/qr.jl:55-62
            if m >= n:
                return UpperTriangular(self.factors[:n, :n])
            raise ArgumentError("R matrix is trapezoid and cannot be extracted with indexing")
        if key is QBlocked:                                     # src/qr.jl:64-83
            k = min(m, n)
            F = self.factors
            T = empty_colmajor(max(k, 1), max(k, 1), dtype=F.dtype, device=F.device)[:k, :k]
            ldf = _mat(F, "F")
            f = getattr(lib(), f"gla_qr_tblock_{_sfx(F)}")
            with torch.cuda.device(F.device):
                rc = f(_handle(F.device.index), m, n, C.c_void_p(F.data_ptr()), ldf,
                       C.c_void_p(self.τ.data_ptr()), C.c_void_p(T.data_ptr()), max(k, 1),
                       _stream(F))
            _check(rc, "gla_qr_tblock")
            return HouseholderBlock(F, UpperTriangular(T))
        raise KeyError(key)


def qrUnblocked_(A, tau=None):
    """qrUnblocked!(A, τ) — src/qr.jl:86-111.  Returns QR2(A, τ)."""
    lda = _mat(A)
    m, n = A.shape
    k = min(m, n)
    if tau is None:
        tau = torch.zeros(max(k, 1), dtype=A.dtype, device=A.device)[:k]
    _vec(tau, k, A.dtype)
    f = getattr(lib(), f"gla_qr_unblocked_{_sfx(A)}")
    with torch.cuda.device(A.device):
        rc = f(_handle(A.device.index), m, n, C.c_void_p(A.data_ptr()), lda,
               C.c_void_p(tau.data_ptr()), _stream(A))
    _check(rc, "gla_qr_unblocked")
    return QR2(A, tau)


def qrBlocked_(A, blocksize: int = 12, tau=None, work=None):
    """qrBlocked!(A, blocksize = 12, τ, work) — src/qr.jl:113-146.  Returns QR2(A, τ)."""
    lda = _mat(A)
    m, n = A.shape
    k = min(m, n)
    if int(blocksize) < 1:
        raise ArgumentError("blocksize must be positive")
    if tau is None:
        tau = torch.zeros(max(k, 1), dtype=A.dtype, device=A.device)[:k]
    _vec(tau, k, A.dtype)
    if work is not None:
        ldw = _mat(work, "work")
        wp = C.c_void_p(work.data_ptr())
    else:
        ldw, wp = max(int(blocksize), 1), C.c_void_p(None)
    f = getattr(lib(), f"gla_qr_blocked_{_sfx(A)}")
    with torch.cuda.device(A.device):
        if A.is_complex():
            if work is not None:
                raise ArgumentError("the complex qrBlocked! keeps its workspace internal")
            rc = f(_handle(A.device.index), m, n, C.c_void_p(A.data_ptr()), lda, int(blocksize),
                   C.c_void_p(tau.data_ptr()), _stream(A))
        else:
            rc = f(_handle(A.device.index), m, n, C.c_void_p(A.data_ptr()), lda, int(blocksize),
                   C.c_void_p(tau.data_ptr()), wp, ldw, _stream(A))
    _check(rc, "gla_qr_blocked")
    return QR2(A, tau)


# ----------------------------------------------------------- batched ----
import json as _json  # noqa: E402
import os as _os  # noqa: E402

_TUNING = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "batched_tuning.json")


def batched_tuning(dtype) -> dict:
    """Launch shape of the batched solver: the autotuned table
    (gla/batched_tuning.json, written by gla.autotune on an MI355X) or
    defaults."""
    key = "float64" if dtype in (torch.float64, np.float64, "float64") else "float32"
    cfg = {"wave_max": 256, "grid_mult": 12}
    if _os.path.exists(_TUNING):
        with open(_TUNING) as f:
            cfg.update(_json.load(f).get(key, {}))
    return cfg


class BatchPlan:
    """Sizes of a batch of SPD matrices sorted into the solver's bins (see
    gla_batch_plan_create).  Build once, reuse for every batched call with
    the same sizes (this is what a captured graph replays)."""

    def __init__(self, sizes, dtype=torch.float64, wave_max=None, grid_mult=None, device=0):
        cfg = batched_tuning(dtype)
        self.wave_max = int(cfg["wave_max"] if wave_max is None else wave_max)
        self.grid_mult = int(cfg["grid_mult"] if grid_mult is None else grid_mult)
        self.sizes = np.ascontiguousarray(np.asarray(sizes, dtype=np.int32))
        self.device = device
        self.dtype = dtype
        hp = C.c_void_p()
        with torch.cuda.device(device):
            _check(lib().gla_batch_plan_create(_handle(device), len(self.sizes),
                                               self.sizes.ctypes.data_as(C.c_void_p),
                                               self.wave_max, self.grid_mult, C.byref(hp)),
                   "gla_batch_plan_create")
        self._plan = hp.value

    def __del__(self):
        try:
            if getattr(self, "_plan", None):
                lib().gla_batch_plan_destroy(self._plan)
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass


def chol_solve_batched_(plan: BatchPlan, A_ptrs, lda, b_ptrs, info=None, stream=None):
    """Factor A_i = L_i L_iᵀ in place and overwrite b_i with A_i⁻¹ b_i for
    every matrix of the plan, in one launch.  A_ptrs / b_ptrs are int64
    device tensors of pointers, lda an int32 device tensor, info an optional
    int32 device tensor (first negative pivot per matrix, 0 if none)."""
    sfx = "d" if plan.dtype == torch.float64 else "s"
    f = getattr(lib(), f"gla_chol_solve_batched_{sfx}")
    st = C.c_void_p(stream if stream is not None else torch.cuda.current_stream().cuda_stream)
    rc = f(_handle(plan.device), C.c_void_p(plan._plan), C.c_void_p(A_ptrs.data_ptr()),
           C.c_void_p(lda.data_ptr()), C.c_void_p(b_ptrs.data_ptr()),
           C.c_void_p(info.data_ptr() if info is not None else None), st)
    _check(rc, "gla_chol_solve_batched")


class PackedBatch:
    """A batch of n_i×n_i matrices and n_i-vectors packed in two device
    buffers (column-major, lda = n_i), with the pointer tables the batched
    entry point takes."""

    def __init__(self, sizes, dtype=torch.float64, device="cuda"):
        self.sizes = np.asarray(sizes, dtype=np.int64)
        self.dtype = dtype
        esz = torch.empty((), dtype=dtype).element_size()
        self.aoff = np.concatenate([[0], np.cumsum(self.sizes * self.sizes)])
        self.boff = np.concatenate([[0], np.cumsum(self.sizes)])
        self.A = torch.zeros(int(self.aoff[-1]) + 1, dtype=dtype, device=device)
        self.b = torch.zeros(int(self.boff[-1]) + 1, dtype=dtype, device=device)
        ab, bb = self.A.data_ptr(), self.b.data_ptr()
        self.A_ptrs = torch.tensor(ab + self.aoff[:-1] * esz, dtype=torch.int64, device=device)
        self.b_ptrs = torch.tensor(bb + self.boff[:-1] * esz, dtype=torch.int64, device=device)
        self.lda = torch.tensor(np.maximum(self.sizes, 1), dtype=torch.int32, device=device)
        self.info = torch.zeros(len(self.sizes), dtype=torch.int32, device=device)

    def matrix(self, i) -> torch.Tensor:
        n = int(self.sizes[i])
        return self.A[int(self.aoff[i]):int(self.aoff[i]) + n * n].view(n, n).t()

    def vector(self, i) -> torch.Tensor:
        n = int(self.sizes[i])
        return self.b[int(self.boff[i]):int(self.boff[i]) + n]


# ------------------------------------------- mixed precision + refinement
class MixedSolveResult:
    """Outcome of chol_solve_mixed: the precision of the returned X, the
    refinement corrections made in it, and the final normwise backward
    error max_k ‖b_k − A x_k‖ / (‖A‖_F ‖x_k‖ + ‖b_k‖)."""

    def __init__(self, iters: int, berr: float, precision: str, tol: float):
        self.iters = iters
        self.berr = berr
        self.precision = precision
        self.converged = berr <= tol

    def __repr__(self):
        return (f"MixedSolveResult(precision={self.precision}, iters={self.iters}, "
                f"berr={self.berr:.3e}, converged={self.converged})")


_PREC = {"fp64": b"d", "fp32": b"s", "bf16": b"B", "fp16": b"H", "bf16x1": b"b",
         "fp16x1": b"h", "auto": b"a"}
_PREC_NAME = {b"d": "fp64", b"s": "fp32", b"B": "bf16", b"H": "fp16", b"b": "bf16x1",
              b"h": "fp16x1"}


def chol_solve_mixed(A, B, precision: str = "auto", tol: float = 1e-12, max_iter: int = 30,
                     X=None):
    """X = A⁻¹ B for SPD Float64 A (lower triangle read), factored in
    `precision` and refined in Float64 to a normwise backward error ≤ tol:
    "fp64"; "fp32"; "bf16": FP32 storage, FP64 diagonal blocks, trailing
    updates on BF16 MFMA with FP32 accumulation from three BF16 terms of
    each panel; "fp16": the same with two scaled FP16 terms; "bf16x1" /
    "fp16x1": one term (faster, for κ up to about 1e3 / 1e4); "auto": fp32,
    then fp64 if fp32 does not converge.  Raises DomainError (index = first
    non-positive pivot) when a fixed-precision factorisation breaks down.
    Returns (X, MixedSolveResult); check .converged for fixed precisions."""
    if precision not in _PREC:
        raise ArgumentError(f"precision must be one of {sorted(_PREC)}")
    if A.dtype != torch.float64 or B.dtype != torch.float64:
        raise TypeError("A and B must be Float64")
    lda = _mat(A)
    vec = B.dim() == 1
    Bm = B.view(-1, 1) if vec else B
    ldb = _mat(Bm, "B")
    n = A.shape[0]
    if A.shape[1] != n:
        raise DimensionMismatch(f"matrix is not square: dimensions are {tuple(A.shape)}")
    if Bm.shape[0] != n:
        raise DimensionMismatch(f"B has {Bm.shape[0]} rows, A is {n}×{n}")
    nrhs = Bm.shape[1]
    if X is None:
        X = empty_colmajor(n, nrhs, torch.float64, device=A.device)
    ldx = _mat(X, "X")
    it = C.c_int32(0)
    be = C.c_double(0.0)
    info = C.c_int64(0)
    used = C.create_string_buffer(1)
    with torch.cuda.device(A.device):
        rc = lib().gla_chol_solve_mixed_d(_handle(A.device.index or 0), _PREC[precision], n,
                                          C.c_void_p(A.data_ptr()), lda, nrhs,
                                          C.c_void_p(Bm.data_ptr()), ldb,
                                          C.c_void_p(X.data_ptr()), ldx, float(tol),
                                          int(max_iter), C.byref(it), C.byref(be),
                                          C.byref(info), used, _stream(A))
    _check(rc, "gla_chol_solve_mixed_d")
    if info.value != 0:
        raise DomainError(info.value)
    res = MixedSolveResult(it.value, be.value, _PREC_NAME.get(used.raw[:1], "?"), tol)
    return (X.view(-1) if vec else X), res


_LOWP_TIN = {"bf16x1": b"b", "fp16x1": b"h", "bf16": b"B", "fp16": b"H"}


def gemm_lowp_nt_(Cm, A, B, alpha=-1.0, precision: str = "bf16x1"):
    """Cm (FP32) += alpha · A Bᵀ on BF16/FP16 MFMA with FP32 accumulation
    (the trailing update of the mixed factorisation): "bf16x1"/"fp16x1" round
    A and B once; "bf16" / "fp16" split them into three BF16 / two scaled
    FP16 terms."""
    if precision not in _LOWP_TIN:
        raise ArgumentError(f"precision must be one of {sorted(_LOWP_TIN)}")
    for X, nm in ((Cm, "C"), (A, "A"), (B, "B")):
        if X.dtype != torch.float32:
            raise TypeError(f"{nm} must be Float32")
    ldc, lda, ldb = _mat(Cm, "C"), _mat(A), _mat(B, "B")
    m, n = Cm.shape
    if A.shape[0] != m or B.shape[0] != n or A.shape[1] != B.shape[1]:
        raise DimensionMismatch("C (m×n) += A (m×k) Bᵀ (k×n)")
    with torch.cuda.device(Cm.device):
        rc = lib().gla_gemm_lowp_nt(_handle(Cm.device.index or 0), _LOWP_TIN[precision], m, n,
                                    A.shape[1], float(alpha), C.c_void_p(A.data_ptr()), lda,
                                    C.c_void_p(B.data_ptr()), ldb, C.c_void_p(Cm.data_ptr()),
                                    ldc, _stream(Cm))
    _check(rc, "gla_gemm_lowp_nt")
    return Cm


# ------------------------------------------------------------------ adjoints
def chol_adjoint(L, Lbar, convention: str = "S", out=None):
    """Ā of A = LLᵀ for the cotangent L̄ (lower parts of L and L̄ read):
    ½ L⁻ᵀ(P + Pᵀ)L⁻¹ with P = Φ(LᵀL̄).  convention "S": symmetric; "L": the
    derivative with respect to the lower triangle (2Ā below the diagonal, Ā
    on it, 0 above)."""
    if convention not in ("S", "L"):
        raise ArgumentError("convention must be 'S' or 'L'")
    ldl, ldg = _mat(L, "L"), _mat(Lbar, "L̄")
    n = L.shape[0]
    if L.shape != (n, n) or Lbar.shape != (n, n):
        raise DimensionMismatch("L and L̄ must be n×n")
    if Lbar.dtype != L.dtype:
        raise TypeError("L and L̄ must have one element type")
    if out is None:
        out = empty_colmajor(n, n, L.dtype, device=L.device)
    lda = _mat(out, "Ā")
    with torch.cuda.device(L.device):
        rc = getattr(lib(), f"gla_chol_adjoint_{_sfx(L)}")(
            _handle(L.device.index), convention.encode(), n, C.c_void_p(L.data_ptr()), ldl,
            C.c_void_p(Lbar.data_ptr()), ldg, C.c_void_p(out.data_ptr()), lda, _stream(L))
    _check(rc, "gla_chol_adjoint")
    return out


def chol_solve_(F, B, uplo=Val.L):
    """ldiv! with a Cholesky factor (LAPACK potrs): B <- A⁻¹ B for A = L Lᵀ
    (uplo Val.L, L in F's lower triangle as cholRecursive! leaves it) or
    A = Uᵀ U (Val.U).  B is n×nrhs (or a vector).  Returns B."""
    if isinstance(F, (LowerTriangular, UpperTriangular)):
        uplo = Val.L if isinstance(F, LowerTriangular) else Val.U
        F = F.data
    ldf = _mat(F, "F")
    n = F.shape[0]
    if F.shape != (n, n):
        raise DimensionMismatch("the factor must be square")
    vec = B.dim() == 1
    Bm = B.reshape(-1, 1) if vec else B
    ldb = _mat(Bm, "B")
    if Bm.shape[0] != n:
        raise DimensionMismatch(f"B has {Bm.shape[0]} rows, the factor is {n}×{n}")
    if Bm.dtype != F.dtype:
        raise TypeError("F and B must have one element type")
    with torch.cuda.device(F.device):
        rc = getattr(lib(), f"gla_chol_solve_{_sfx(F)}")(
            _handle(F.device.index), _uplo(uplo).encode(), n, Bm.shape[1],
            C.c_void_p(F.data_ptr()), ldf, C.c_void_p(Bm.data_ptr()), ldb, _stream(F))
    _check(rc, "gla_chol_solve")
    return B


def qr_solve_(F, tau, B=None):
    """ldiv! with a QR (least squares, m ≥ n): B (m×nrhs, or a vector)
    <- Qᵀ B with its first n rows then replaced by R⁻¹ of themselves, the
    minimiser of ‖A X − B‖.  F, τ as qrBlocked!/qrUnblocked! leave them (or
    a QR2 as the first argument).  Returns X = B[:n].  Raises
    ArgumentError("R is singular ...") when R has a zero diagonal entry."""
    if isinstance(F, QR2):
        F, tau, B = F.factors, F.tau, tau
    ldf = _mat(F, "F")
    m, n = F.shape
    if n > m:
        raise ArgumentError("least squares through QR needs m >= n")
    _vec(tau, n, F.dtype)
    vec = B.dim() == 1
    Bm = B.reshape(-1, 1) if vec else B
    ldb = _mat(Bm, "B")
    if Bm.shape[0] != m:
        raise DimensionMismatch(f"B has {Bm.shape[0]} rows, A has {m}")
    if Bm.dtype != F.dtype:
        raise TypeError("F and B must have one element type")
    info = C.c_int64(0)
    with torch.cuda.device(F.device):
        rc = getattr(lib(), f"gla_qr_solve_{_sfx(F)}")(
            _handle(F.device.index), m, n, Bm.shape[1], C.c_void_p(F.data_ptr()), ldf,
            C.c_void_p(tau.data_ptr()), C.c_void_p(Bm.data_ptr()), ldb, C.byref(info), _stream(F))
    _check(rc, "gla_qr_solve")
    if info.value > 0:
        raise ArgumentError(f"R is singular: zero diagonal entry at {info.value}")
    return B[:n]


# ------------------------------------------------------------------ LDLᵀ ----
class LDLt:
    """``LDLt(Hermitian(A, uplo))`` as ldlt! returns it (src/ldlt.jl:150-157):
    ``data`` holds the factors in the ``uplo`` triangle.  Properties ``d``
    (the diagonal of D, a view), ``L`` (unit lower as a dense tensor; Uᵀ for
    uplo 'U'), and ``ldiv_``/``lmul_``/``@`` as src/ldlt.jl:225-234 define
    them."""

    def __init__(self, data, uplo):
        self.data = data
        self.uplo = uplo

    @property
    def d(self):
        return torch.diagonal(self.data)

    @property
    def L(self):
        n = self.data.shape[0]
        E = torch.eye(n, dtype=self.data.dtype, device=self.data.device)
        if self.uplo == "L":
            return torch.tril(self.data, -1) + E
        return torch.triu(self.data, 1).t() + E

    def _apply(self, X, fn):
        F = self.data
        ldf = _mat(F, "F")
        n = F.shape[0]
        vec = X.dim() == 1
        Xm = X.reshape(-1, 1) if vec else X
        ldx = _mat(Xm, "X")
        if Xm.shape[0] != n:
            raise DimensionMismatch(f"X has {Xm.shape[0]} rows, the factorisation is {n}×{n}")
        if Xm.dtype != F.dtype:
            raise TypeError("F and X must have one element type")
        with torch.cuda.device(F.device):
            rc = getattr(lib(), f"{fn}_{_sfx(F)}")(
                _handle(F.device.index), self.uplo.encode(), n, Xm.shape[1],
                C.c_void_p(F.data_ptr()), ldf, C.c_void_p(Xm.data_ptr()), ldx, _stream(F))
        _check(rc, fn)
        return X

    def ldiv_(self, X):
        """ldiv!(F, X) — src/ldlt.jl:225-228: X <- A⁻¹ X in place."""
        return self._apply(X, "gla_ldlt_solve")

    def lmul_(self, X):
        """lmul!(F, X) — src/ldlt.jl:229-232: X <- A X in place."""
        return self._apply(X, "gla_ldlt_lmul")

    def __matmul__(self, X):
        """F * X — src/ldlt.jl:234 (a copy of X, then lmul!)."""
        return self.lmul_(colmajor(X, device=X.device))


def ldlt_(Ah, blocksize=None):
    """ldlt!(A::Hermitian, blocksize = 128 ÷ sizeof(T)) — src/ldlt.jl:150-157.
    Factors Ah.data in place (only its ``uplo`` triangle is read and
    written); returns LDLt."""
    if not isinstance(Ah, Hermitian):
        raise TypeError("ldlt! takes a Hermitian")
    A = Ah.data
    lda = _mat(A)
    n, m = A.shape
    if n != m:
        raise DimensionMismatch(f"matrix is not square: dimensions are {tuple(A.shape)}")
    if blocksize is None:
        blocksize = max(1, 128 // A.element_size())
    with torch.cuda.device(A.device):
        rc = getattr(lib(), f"gla_ldlt_{_sfx(A)}")(
            _handle(A.device.index), Ah.uplo.encode(), n, C.c_void_p(A.data_ptr()), lda,
            int(blocksize), _stream(A))
    _check(rc, "gla_ldlt")
    return LDLt(A, Ah.uplo)


def ldlt(Ah, blocksize=None):
    """ldlt(A::Hermitian) — src/ldlt.jl:202: ldlt! of a copy."""
    return ldlt_(Hermitian(colmajor(Ah.data, device=Ah.data.device), Ah.uplo), blocksize)


# ------------------------------------------------- two-sided reductions ----
class SymmetricTridiagonalFactorization:
    """What symtri! returns (src/eigenSelfAdjoint.jl:450-564): the reflectors
    (``factors`` with ``uplo`` and ``tau``, the EigenQ) and the tridiagonal
    (``diagonals`` = (d, e))."""

    def __init__(self, uplo, factors, tau):
        self.uplo = uplo
        self.factors = factors
        self.tau = tau

    @property
    def diagonals(self):
        d = torch.diagonal(self.factors).clone()
        e = torch.diagonal(self.factors, -1 if self.uplo == "L" else 1).clone()
        return d, e


def symtri_(Ah, tau=None):
    """symtri!(A::Hermitian) — symtriLower!/symtriUpper!,
    src/eigenSelfAdjoint.jl:446-564, for real element types.  Reduces
    Ah.data in place; returns SymmetricTridiagonalFactorization."""
    if not isinstance(Ah, Hermitian):
        raise TypeError("symtri! takes a Hermitian (or Symmetric) matrix")
    A = Ah.data
    lda = _mat(A)
    n, m = A.shape
    if n != m:
        raise DimensionMismatch(f"matrix is not square: dimensions are {tuple(A.shape)}")
    if tau is None:
        tau = torch.zeros(max(n - 1, 0), dtype=A.dtype, device=A.device)
    _vec(tau, max(n - 1, 0), A.dtype)
    with torch.cuda.device(A.device):
        rc = getattr(lib(), f"gla_symtri_{_sfx(A)}")(
            _handle(A.device.index), Ah.uplo.encode(), n, C.c_void_p(A.data_ptr()), lda,
            C.c_void_p(tau.data_ptr()), _stream(A))
    _check(rc, "gla_symtri")
    return SymmetricTridiagonalFactorization(Ah.uplo, A, tau)


class BidiagonalFactorization:
    """What bidiagonalize! returns (src/svd.jl:321-381): the bidiagonal
    (``uplo`` 'U' for m ≥ n, 'L' otherwise; ``dv``, ``ev``), the reflectors
    (``factors``) and their scalars ``taul``, ``taur``."""

    def __init__(self, factors, taul, taur):
        self.factors = factors
        self.taul = taul
        self.taur = taur
        m, n = factors.shape
        self.uplo = "U" if m >= n else "L"

    @property
    def bidiagonal(self):
        F = self.factors
        k = min(F.shape)
        dv = torch.diagonal(F)[:k].clone()
        ev = torch.diagonal(F, 1 if self.uplo == "U" else -1)[:max(k - 1, 0)].clone()
        return dv, ev


def bidiagonalize_(A):
    """bidiagonalize!(A) — src/svd.jl:328-381, real element types.  Reduces A
    in place; returns BidiagonalFactorization."""
    lda = _mat(A)
    m, n = A.shape
    k = min(m, n)
    nl, nr = (k, max(k - 1, 0)) if m >= n else (max(k - 1, 0), k)
    taul = torch.zeros(max(nl, 1), dtype=A.dtype, device=A.device)
    taur = torch.zeros(max(nr, 1), dtype=A.dtype, device=A.device)
    with torch.cuda.device(A.device):
        rc = getattr(lib(), f"gla_bidiag_{_sfx(A)}")(
            _handle(A.device.index), m, n, C.c_void_p(A.data_ptr()), lda,
            C.c_void_p(taul.data_ptr()), C.c_void_p(taur.data_ptr()), _stream(A))
    _check(rc, "gla_bidiag")
    return BidiagonalFactorization(A, taul[:nl], taur[:nr])


def _batch3(X, name):
    """(batch, n, n) device tensor of column-major matrices: returns (ld, stride)."""
    if not isinstance(X, torch.Tensor) or X.dim() != 3 or not X.is_cuda:
        raise TypeError(f"{name} must be a 3-D device tensor (batch, n, n)")
    b, m, n = X.shape
    if m > 1 and X.stride(1) != 1:
        raise TypeError(f"{name}: each matrix must be column-major (stride(1) == 1)")
    return max(X.stride(2) if n > 1 else m, 1), (X.stride(0) if b > 1 else m * n)


def chol_adjoint_batched(A, Lbar, convention: str = "S", out=None):
    """Ā for a batch of SPD matrices A[b] (order n ≤ 128, lower triangles
    read) and cotangents L̄[b], recomputing each factor L[b] = chol(A[b]):
    the chol_adjoint of every matrix, in one pass of batched launches.
    Tensors are (batch, n, n) with column-major matrices.  Raises
    DomainError naming the first matrix whose factorisation breaks down."""
    if convention not in ("S", "L"):
        raise ArgumentError("convention must be 'S' or 'L'")
    lda, sa = _batch3(A, "A")
    ldg, sg = _batch3(Lbar, "L̄")
    b, n, n2 = A.shape
    if n != n2 or Lbar.shape != A.shape:
        raise DimensionMismatch("A and L̄ must be (batch, n, n)")
    if Lbar.dtype != A.dtype:
        raise TypeError("A and L̄ must have one element type")
    if out is None:
        out = torch.empty((b, n, n), dtype=A.dtype, device=A.device).transpose(1, 2)
    ldo, so = _batch3(out, "Ā")
    info = torch.zeros(max(b, 1), dtype=torch.int32, device=A.device)
    with torch.cuda.device(A.device):
        rc = getattr(lib(), f"gla_chol_adjoint_batched_{_sfx(A)}")(
            _handle(A.device.index), convention.encode(), n, b, C.c_void_p(A.data_ptr()), lda, sa,
            C.c_void_p(Lbar.data_ptr()), ldg, sg, C.c_void_p(out.data_ptr()), ldo, so,
            C.c_void_p(info.data_ptr()), _stream(A))
    _check(rc, "gla_chol_adjoint_batched")
    bad = torch.nonzero(info[:b]).flatten()
    if bad.numel():
        k = int(bad[0])
        raise DomainError(f"matrix {k}: sqrt of a negative pivot at position {int(info[k])}")
    return out


def qr_adjoint(F, tau, Rbar, out=None):
    """Ā (m×n, m ≥ n) of A = QR stored as (F, τ) by qrBlocked!/qrUnblocked!,
    for the cotangent R̄ (upper part read) and Q̄ = 0:
    Q [copyltu(R R̄ᵀ) R⁻ᵀ ; 0]."""
    ldf, ldr = _mat(F, "F"), _mat(Rbar, "R̄")
    m, n = F.shape
    if n > m:
        raise ArgumentError("the QR adjoint needs m >= n")
    if Rbar.shape != (n, n):
        raise DimensionMismatch("R̄ must be n×n")
    _vec(tau, n, F.dtype)
    if out is None:
        out = empty_colmajor(m, n, F.dtype, device=F.device)
    lda = _mat(out, "Ā")
    with torch.cuda.device(F.device):
        rc = getattr(lib(), f"gla_qr_adjoint_{_sfx(F)}")(
            _handle(F.device.index), m, n, C.c_void_p(F.data_ptr()), ldf,
            C.c_void_p(tau.data_ptr()), C.c_void_p(Rbar.data_ptr()), ldr,
            C.c_void_p(out.data_ptr()), lda, _stream(F))
    _check(rc, "gla_qr_adjoint")
    return out


class CholeskyFn(torch.autograd.Function):
    """torch.autograd rule for L = cholRecursive!(copy(A)) (lower part of A
    read; the returned L has zeros above the diagonal).  The backward is
    gla_chol_adjoint with the symmetric convention, the analogue of a
    ChainRulesCore rrule for the Julia binding."""

    @staticmethod
    def forward(ctx, A):
        W = colmajor(A)
        cholRecursive_(W, Val.L, 1)
        L = torch.tril(W)
        ctx.save_for_backward(L)
        return L

    @staticmethod
    def backward(ctx, gL):
        (L,) = ctx.saved_tensors
        Lc = colmajor(L)
        return chol_adjoint(Lc, colmajor(gL, L.dtype), "S").contiguous()


class QRFn(torch.autograd.Function):
    """torch.autograd rule for R of qrBlocked!(copy(A), blocksize) (m ≥ n,
    R̄ only): backward by gla_qr_adjoint."""

    @staticmethod
    def forward(ctx, A, blocksize=128):
        q = qrBlocked_(colmajor(A), blocksize)
        F, tau = q.factors, q.tau
        ctx.save_for_backward(F, tau)
        return torch.triu(F[: F.shape[1], :].contiguous())

    @staticmethod
    def backward(ctx, gR):
        F, tau = ctx.saved_tensors
        return qr_adjoint(F, tau, colmajor(gR, F.dtype)).contiguous(), None

