"""tmv_amd -- MI355X (gfx950) dense GEMM + LU behind TMV's interface.

The product is ``tmv_amd/libtmvx.so`` (HIP kernels + the C ABI declared in
``include/tmvx.h``).  This module is the Python host mirror of that ABI:
thin ctypes bindings taking torch device tensors (torch is used only for
device memory and streams).  There is no CPU fallback: every call goes to
the HIP library and fails loudly when it is missing.

Matrix convention: a column-major m x n matrix is a torch tensor of shape
(m, n) with stride (1, ld), e.g. ``colmajor(t)`` or ``torch.empty(n, m).t()``;
a row-major one has stride (ld, 1).
"""
from __future__ import annotations

import ctypes
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtmvx.so")

OK = 0
ERR_ARG = -1
ERR_ALLOC = -2
ERR_HIP = -3
ERR_SINGULAR = -4
ERR_NOT_CONVERGED = -5
ERR_TIMEOUT = -6
ERR_UNSUPPORTED = -7

COLMAJOR = 0
ROWMAJOR = 1

_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_d = ctypes.c_double
_f = ctypes.c_float

# name -> (restype, argtypes)
_SIGS = {
    "tmvx_status_str": (ctypes.c_char_p, [_int]),
    "tmvx_version": (ctypes.c_char_p, []),
    "tmvx_gemm_f64": (_int, [_int, _int, _i64, _i64, _i64, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64, _vp]),
    "tmvx_gemm_f32": (_int, [_int, _int, _i64, _i64, _i64, _f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _vp]),
    "tmvx_getrf_f64": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrf_f64_noalloc": (_int, [_i64, _vp, _i64, _vp, _vp, _vp]),
    "tmvx_getrf_f32": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrs_f64": (_int, [_int, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _int, _vp]),
    "tmvx_getrs_f32": (_int, [_int, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _int, _vp]),
    "tmvx_gemm_z": (_int, [_int, _int, _int, _int, _i64, _i64, _i64, _d, _d, _vp, _i64, _vp, _i64, _d, _d, _vp, _i64, _vp]),
    "tmvx_gemm_c": (_int, [_int, _int, _int, _int, _i64, _i64, _i64, _f, _f, _vp, _i64, _vp, _i64, _f, _f, _vp, _i64, _vp]),
    "tmvx_getrf_z": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrf_c": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrs_z": (_int, [_int, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_getrs_c": (_int, [_int, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_geqrf_f64": (_int, [_i64, _i64, _vp, _i64, _vp, _vp]),
    "tmvx_geqrf_f32": (_int, [_i64, _i64, _vp, _i64, _vp, _vp]),
    "tmvx_geqrs_f64": (_int, [_i64, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_geqrs_f32": (_int, [_i64, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_logdet_z": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_getri_z": (_int, [_i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_getri_f64": (_int, [_i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_getri_f32": (_int, [_i64, _vp, _i64, _vp, _vp, _i64, _vp]),
    "tmvx_logdet_f64": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_logdet_f32": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_trsm_f64": (_int, [_int, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "tmvx_trsm_f32": (_int, [_int, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "tmvx_trsm_blockinv_f64": (_int, [_int, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "tmvx_trsm_blockinv_f32": (_int, [_int, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "tmvx_laswp_f64": (_int, [_int, _i64, _vp, _i64, _i64, _i64, _vp, _vp]),
    "tmvx_laswp_f32": (_int, [_int, _i64, _vp, _i64, _i64, _i64, _vp, _vp]),
    "tmvx_fill_uniform_f64": (_int, [ctypes.c_uint64, ctypes.c_uint64, _i64, _d, _d, _vp, _vp]),
    "tmvx_fill_normal_f64": (_int, [ctypes.c_uint64, ctypes.c_uint64, _i64, _vp, _vp]),
    "tmvx_gesv_mixed_f64": (_int, [_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _d, _int, _vp, _vp]),
    # batched systems (include/tmvx_batched.h)
    "tmvx_batch_plan_create": (_int, [_int, _i64, _vp, _vp, _int, ctypes.POINTER(_vp), _vp]),
    "tmvx_batch_plan_destroy": (_int, [_vp]),
    "tmvx_batch_plan_describe": (ctypes.c_char_p, [_vp]),
    "tmvx_batch_plan_set": (_int, [_vp, _int, _int, _int, _vp]),
    "tmvx_batch_schedule_count": (_int, []),
    "tmvx_gesv_batched_f64": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tmvx_gesv_batched_f32": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tmvx_getrs_batched_f64": (_int, [_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tmvx_getrs_batched_f32": (_int, [_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tmvx_gemm_batched_f64": (_int, [_i64, _vp, _vp, _vp, _d, _vp, _vp, _d, _vp, _vp]),
    "tmvx_gemm_batched_f32": (_int, [_i64, _vp, _vp, _vp, _f, _vp, _vp, _f, _vp, _vp]),
    "tmvx_getrf_f32_tournament": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrf_nopiv_f64": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrf_nopiv_f32": (_int, [_i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_lu_diag_range_f64": (_int, [_i64, _vp, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_lu_diag_range_f32": (_int, [_i64, _vp, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_lu_diag_range_z": (_int, [_i64, _vp, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_lu_diag_range_c": (_int, [_i64, _vp, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d), _vp]),
    "tmvx_getrf_mn_f64": (_int, [_i64, _i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_getrf_mn_f32": (_int, [_i64, _i64, _vp, _i64, _vp, ctypes.POINTER(_i64), _vp]),
    "tmvx_panel_timer_start": (_int, [_int]),
    "tmvx_panel_timer_stop": (_int, [ctypes.POINTER(_i64), ctypes.POINTER(_d)]),
    "tmvx_gemm_f32_lowp": (_int, [_int, _i64, _i64, _i64, _f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _int, _vp]),
    "tmvx_getrf_batched_f64": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "tmvx_getrf_batched_f32": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "tmvx_gesv_adjoint_f64": (_int, [_i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int, _vp]),
    "tmvx_gesv_adjoint_f32": (_int, [_i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int, _vp]),
    "tmvx_gesv_adjoint_batched_f64": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    "tmvx_gesv_adjoint_batched_f32": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    "tmvx_gesv_adjoint_batched_recompute_f64": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    "tmvx_gesv_adjoint_batched_recompute_f32": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    # tuning hooks (include/tmvx_tune.h)
    "tmvx_gemm_config_count": (_int, [_int]),
    "tmvx_gemm_config_name": (ctypes.c_char_p, [_int, _int]),
    "tmvx_gemm_cfg_f64": (_int, [_int, _int, _int, _i64, _i64, _i64, _d, _vp, _i64, _vp, _i64, _d, _vp, _i64, _vp]),
    "tmvx_gemm_cfg_f32": (_int, [_int, _int, _int, _i64, _i64, _i64, _f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _vp]),
}

_lib = None


class TmvxError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {status_str(status)} ({status})")


class SingularError(TmvxError):
    """A zero diagonal met a non-zero right-hand side (TMV's Singular)."""


def lib() -> ctypes.CDLL:
    """The HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m tmv_amd.build`"
                " (there is no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGS)


def status_str(status: int) -> str:
    return lib().tmvx_status_str(status).decode()


def _check(rc: int, where: str) -> None:
    if rc == ERR_SINGULAR:
        raise SingularError(rc, where)
    if rc != OK:
        raise TmvxError(rc, where)


# ---------------------------------------------------------------------------
# torch plumbing
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _stream(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def colmajor(t):
    """Column-major copy of a 2-D tensor (stride (1, m))."""
    return t.t().contiguous().t()


def empty_colmajor(m, n, dtype, device="cuda"):
    torch = _torch()
    return torch.empty((n, m), dtype=dtype, device=device).t()


def _layout(t, name: str):
    """(layout, ld) of a 2-D tensor that is column- or row-major."""
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D")
    m, n = t.shape
    s0, s1 = t.stride()
    if m == 0 or n == 0:                       # no elements: any layout
        return COLMAJOR, max(1, m)
    if s0 == 1 and (n <= 1 or s1 >= max(1, m)):
        return COLMAJOR, max(s1, m, 1)
    if s1 == 1 and (m <= 1 or s0 >= max(1, n)):
        return ROWMAJOR, max(s0, n, 1)
    raise ValueError(f"{name} must be column- or row-major with unit stride")


def _colmajor_ld(t, name: str) -> int:
    lay, ld = _layout(t, name)
    if lay != COLMAJOR:
        raise ValueError(f"{name} must be column-major")
    return ld


def _on_gpu(*ts) -> None:
    """Every tensor a device tensor, all on one device (the library takes
    device pointers; a host pointer would be read by the GPU)."""
    d = ts[0].device
    for t in ts:
        if t.device.type != "cuda":
            raise ValueError(f"device tensors only (got a {t.device.type} tensor)")
        if t.device != d:
            raise ValueError(f"tensors on different devices ({d} and {t.device})")


def _square(t, name: str) -> int:
    if t.dim() != 2 or t.shape[0] != t.shape[1]:
        raise ValueError(f"{name} must be a square 2-D tensor (got shape {tuple(t.shape)})")
    return t.shape[0]


def _check_piv(piv, n: int) -> None:
    torch = _torch()
    if piv.dtype != torch.int64:
        raise TypeError(f"piv must be int64 (got {piv.dtype})")
    if piv.dim() != 1 or piv.numel() < n or (n > 0 and piv.stride(0) != 1):
        raise ValueError(f"piv must be a contiguous 1-D tensor of at least {n} entries")


def _rhs(B, n: int, dtype, name: str = "B"):
    """(B as n x nrhs, its leading dimension) for a 1-D (one column) or
    column-major 2-D right-hand side of n rows and dtype `dtype`."""
    if B.dim() not in (1, 2):
        raise ValueError(f"{name} must be 1-D or 2-D")
    if B.shape[0] != n:
        raise ValueError(f"{name} has {B.shape[0]} rows, expected {n}")
    if B.dtype != dtype:
        raise TypeError(f"{name} is {B.dtype}, expected {dtype}")
    Bm = B.view(n, 1) if B.dim() == 1 else B
    if Bm.shape[1] > 1:
        return Bm, _colmajor_ld(Bm, name)
    if n > 1 and Bm.stride(0) != 1:
        raise ValueError(f"{name} must have unit row stride")
    return Bm, max(n, 1)


def _suffix(dtype) -> str:
    torch = _torch()
    if dtype == torch.float64:
        return "f64"
    if dtype == torch.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {dtype}")


# ---------------------------------------------------------------------------
# The operations (mirror TMV: MultMM, LU_Decompose, LU_LDivEq / LU_RDivEq)
# ---------------------------------------------------------------------------

def gemm(A, B, C=None, alpha=1.0, beta=0.0, stream=None):
    """C := alpha*A*B + beta*C (TMV MultMM<add> with beta = add).  A and B
    may be column- or row-major; C (allocated column-major when None) must
    be column-major.  Returns C."""
    if A.dim() != 2 or B.dim() != 2:
        raise ValueError("A and B must be 2-D")
    m, k = A.shape
    k2, n = B.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    suf = _suffix(A.dtype)
    if B.dtype != A.dtype or (C is not None and C.dtype != A.dtype):
        raise TypeError("A, B and C must share a dtype")
    if C is not None and tuple(C.shape) != (m, n):
        raise ValueError(f"C has shape {tuple(C.shape)}, expected {(m, n)}")
    la, lda = _layout(A, "A")
    lb, ldb = _layout(B, "B")
    _on_gpu(A, B, *(() if C is None else (C,)))
    if C is None:
        C = empty_colmajor(m, n, A.dtype, A.device)
        beta = 0.0
    ldc = _colmajor_ld(C, "C")
    fn = getattr(lib(), "tmvx_gemm_" + suf)
    _check(fn(la, lb, m, n, k, alpha, A.data_ptr(), lda, B.data_ptr(), ldb, beta,
              C.data_ptr(), ldc, _stream(stream)), "gemm")
    return C


def gemm_complex(A, B, C=None, alpha=1.0, beta=0.0, conj_a=False, conj_b=False, stream=None):
    """C := alpha*op(A)*op(B) + beta*C for complex128 / complex64 tensors
    (TMV MultMM<add> on complex matrices; op = conjugate when conj_x).  A
    and B may be column- or row-major; C (allocated column-major when None)
    must be column-major.  Returns C."""
    torch = _torch()
    m, k = A.shape
    k2, n = B.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    if A.dtype not in (torch.complex128, torch.complex64) or B.dtype != A.dtype:
        raise TypeError("A and B must share a complex dtype")
    la, lda = _layout(A, "A")
    lb, ldb = _layout(B, "B")
    if C is None:
        C = empty_colmajor(m, n, A.dtype, A.device)
        beta = 0.0
    if C.dtype != A.dtype:
        raise TypeError("C must have A's dtype")
    if tuple(C.shape) != (m, n):
        raise ValueError(f"C has shape {tuple(C.shape)}, expected {(m, n)}")
    _on_gpu(A, B, C)
    ldc = _colmajor_ld(C, "C")
    alpha, beta = complex(alpha), complex(beta)
    fn = lib().tmvx_gemm_z if A.dtype == torch.complex128 else lib().tmvx_gemm_c
    _check(fn(la, lb, int(bool(conj_a)), int(bool(conj_b)), m, n, k, alpha.real, alpha.imag,
              A.data_ptr(), lda, B.data_ptr(), ldb, beta.real, beta.imag, C.data_ptr(), ldc,
              _stream(stream)), "gemm_complex")
    return C


def getrf(A, stream=None):
    """In-place PA = LU of a square column-major tensor (TMV LU_Decompose).
    Returns (piv, info): piv an int64 device tensor of 0-based interchanges,
    info the 1-based index of the first zero pivot (0 if none)."""
    torch = _torch()
    n = _square(A, "A")
    suf = _suffix(A.dtype)
    lda = _colmajor_ld(A, "A")
    _on_gpu(A)
    piv = torch.empty(n, dtype=torch.int64, device=A.device)
    info = _i64(0)
    fn = getattr(lib(), "tmvx_getrf_" + suf)
    _check(fn(n, A.data_ptr(), lda, piv.data_ptr(), ctypes.byref(info), _stream(stream)),
           "getrf")
    return piv, int(info.value)


def getrf_mn(A, stream=None):
    """In-place PA = LU of an m x n column-major tensor (TMV LU_Decompose on
    a non-square view).  Returns (piv, info): min(m, n) 0-based interchanges
    and the 1-based first zero pivot (0 if none)."""
    torch = _torch()
    if A.dim() != 2:
        raise ValueError(f"A must be a 2-D tensor (got shape {tuple(A.shape)})")
    m, n = A.shape
    suf = _suffix(A.dtype)
    lda = _colmajor_ld(A, "A")
    _on_gpu(A)
    piv = torch.empty(min(m, n), dtype=torch.int64, device=A.device)
    info = _i64(0)
    fn = getattr(lib(), "tmvx_getrf_mn_" + suf)
    _check(fn(m, n, A.data_ptr(), lda, piv.data_ptr(), ctypes.byref(info), _stream(stream)),
           "getrf_mn")
    return piv, int(info.value)


def getrf_complex(A, stream=None):
    """In-place PA = LU of a square column-major complex128 / complex64
    tensor (TMV LU_Decompose for complex T).  Returns (piv, info)."""
    torch = _torch()
    n = _square(A, "A")
    if A.dtype not in (torch.complex128, torch.complex64):
        raise TypeError("complex128 or complex64 only")
    lda = _colmajor_ld(A, "A")
    _on_gpu(A)
    piv = torch.empty(n, dtype=torch.int64, device=A.device)
    info = _i64(0)
    fn = lib().tmvx_getrf_z if A.dtype == torch.complex128 else lib().tmvx_getrf_c
    _check(fn(n, A.data_ptr(), lda, piv.data_ptr(), ctypes.byref(info), _stream(stream)),
           "getrf_complex")
    return piv, int(info.value)


def getrs_complex(LU, piv, B, trans=False, stream=None):
    """In-place complex solve with the factors: B := A^-1 B (trans: A^-T B)."""
    torch = _torch()
    n = _square(LU, "LU")
    if LU.dtype not in (torch.complex128, torch.complex64):
        raise TypeError("complex128 or complex64 only")
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    Bm, ldb = _rhs(B, n, LU.dtype)
    _on_gpu(LU, piv, B)
    fn = lib().tmvx_getrs_z if LU.dtype == torch.complex128 else lib().tmvx_getrs_c
    _check(fn(int(bool(trans)), n, Bm.shape[1], LU.data_ptr(), lda, piv.data_ptr(),
              Bm.data_ptr(), ldb, _stream(stream)), "getrs_complex")
    return B


def geqrf(A, stream=None):
    """In-place QR (Householder) of a column-major m x n tensor, m >= n
    (TMV QR_Decompose).  Returns beta (device, n)."""
    torch = _torch()
    m, n = A.shape
    lda = _colmajor_ld(A, "A") if n > 1 else max(m, 1)
    beta = torch.empty(n, dtype=A.dtype, device=A.device)
    fn = getattr(lib(), "tmvx_geqrf_" + _suffix(A.dtype))
    _check(fn(m, n, A.data_ptr(), lda, beta.data_ptr(), _stream(stream)), "geqrf")
    return beta


def geqrs(QR, beta, B, stream=None):
    """QR division: B (m x nrhs) := Q^T B, then its first n rows := R^-1 of
    themselves (least squares for m > n).  Returns B."""
    m, n = QR.shape
    suf = _suffix(QR.dtype)
    lda = _colmajor_ld(QR, "QR") if n > 1 else max(m, 1)
    if beta.dtype != QR.dtype or beta.dim() != 1 or beta.numel() < n:
        raise ValueError(f"beta must be a 1-D {QR.dtype} tensor of at least {n} entries")
    Bm, ldb = _rhs(B, m, QR.dtype)
    _on_gpu(QR, beta, B)
    fn = getattr(lib(), "tmvx_geqrs_" + suf)
    _check(fn(m, n, Bm.shape[1], QR.data_ptr(), lda, beta.data_ptr(), Bm.data_ptr(), ldb,
              _stream(stream)), "geqrs")
    return B


def logdet_complex(LU, piv, stream=None):
    """(log|det A|, complex sign) from complex128 factors (TMV LUDiv::logDet
    for complex T); sign 0 and logdet -inf for a singular matrix."""
    n = _square(LU, "LU")
    if LU.dtype != _torch().complex128:
        raise TypeError("complex128 only")
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    _on_gpu(LU, piv)
    ld, sr, si = _d(0.0), _d(0.0), _d(0.0)
    _check(lib().tmvx_logdet_z(n, LU.data_ptr(), lda, piv.data_ptr(), ctypes.byref(ld),
                               ctypes.byref(sr), ctypes.byref(si), _stream(stream)),
           "logdet_complex")
    return ld.value, complex(sr.value, si.value)


def getri_complex(LU, piv, stream=None):
    """A^-1 from complex128 factors; returns a new column-major tensor."""
    torch = _torch()
    n = _square(LU, "LU")
    if LU.dtype != torch.complex128:
        raise TypeError("complex128 only")
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    _on_gpu(LU, piv)
    X = colmajor(torch.empty(n, n, dtype=LU.dtype, device=LU.device))
    _check(lib().tmvx_getri_z(n, LU.data_ptr(), lda, piv.data_ptr(), X.data_ptr(), max(n, 1),
                              _stream(stream)), "getri_complex")
    return X


def getrf_nopiv(A, stream=None):
    """A = LU without row interchanges (the mixed solver's factorisation
    after its random butterfly transformation).  Returns (piv, info): piv is
    the identity list, info the 1-based first zero or non-finite pivot (the
    factors are then unusable) or 0."""
    torch = _torch()
    n = _square(A, "A")
    suf = {torch.float64: "f64", torch.float32: "f32"}.get(A.dtype)
    if suf is None:
        raise TypeError("float64 or float32")
    lda = _colmajor_ld(A, "A")
    _on_gpu(A)
    piv = torch.empty(n, dtype=torch.int64, device=A.device)
    info = _i64(0)
    _check(getattr(lib(), "tmvx_getrf_nopiv_" + suf)(n, A.data_ptr(), lda, piv.data_ptr(),
                                                     ctypes.byref(info), _stream(stream)),
           "getrf_nopiv")
    return piv, int(info.value)


def getrf_tournament(A, stream=None):
    """FP32 PA = LU with tournament pivoting (the mixed solver's panel
    factorisation).  Returns (piv, info) like getrf."""
    torch = _torch()
    n = _square(A, "A")
    if A.dtype != torch.float32:
        raise TypeError("float32 only")
    lda = _colmajor_ld(A, "A")
    _on_gpu(A)
    piv = torch.empty(n, dtype=torch.int64, device=A.device)
    info = _i64(0)
    _check(lib().tmvx_getrf_f32_tournament(n, A.data_ptr(), lda, piv.data_ptr(),
                                           ctypes.byref(info), _stream(stream)), "getrf_tournament")
    return piv, int(info.value)


def panel_timer_start(cap=4096):
    """Record HIP events around the next `cap` panel-kernel launches of the
    factorisations (on the stream each panel runs on)."""
    _check(lib().tmvx_panel_timer_start(int(cap)), "panel_timer_start")


def panel_timer_stop():
    """Stop the panel timer: (launches, summed launch time in ms)."""
    cnt = _i64(0)
    ms = _d(0.0)
    _check(lib().tmvx_panel_timer_stop(ctypes.byref(cnt), ctypes.byref(ms)), "panel_timer_stop")
    return int(cnt.value), float(ms.value)


def getrs(LU, piv, B, trans=False, check_singular=True, stream=None):
    """In-place solve with the factors: B := A^-1 B (trans: A^-T B).
    B is column-major n x nrhs (a 1-D tensor is treated as n x 1)."""
    n = _square(LU, "LU")
    suf = _suffix(LU.dtype)
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    Bm, ldb = _rhs(B, n, LU.dtype)
    _on_gpu(LU, piv, B)
    fn = getattr(lib(), "tmvx_getrs_" + suf)
    _check(fn(int(bool(trans)), n, Bm.shape[1], LU.data_ptr(), lda, piv.data_ptr(),
              Bm.data_ptr(), ldb, int(bool(check_singular)), _stream(stream)), "getrs")
    return B


def getri(LU, piv, stream=None):
    """A^-1 from the factors (TMV LU_Inverse / m.inverse() with LU
    division); returns a new column-major n x n tensor."""
    torch = _torch()
    n = _square(LU, "LU")
    suf = _suffix(LU.dtype)
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    _on_gpu(LU, piv)
    X = colmajor(torch.empty(n, n, dtype=LU.dtype, device=LU.device))
    fn = getattr(lib(), "tmvx_getri_" + suf)
    _check(fn(n, LU.data_ptr(), lda, piv.data_ptr(), X.data_ptr(), max(n, 1), _stream(stream)),
           "getri")
    return X


def lu_diag_range(LU, stream=None):
    """(min_i |U_ii|, max_i |U_ii|) of packed factors, |z| = |re| + |im| for
    complex (the numbers TMV's LUDiv::isSingular compares)."""
    torch = _torch()
    n = _square(LU, "LU")
    lda = _colmajor_ld(LU, "LU")
    _on_gpu(LU)
    suf = {torch.float64: "f64", torch.float32: "f32", torch.complex128: "z",
           torch.complex64: "c"}.get(LU.dtype)
    if suf is None:
        raise TypeError("float64, float32, complex128 or complex64 only")
    mn, mx = _d(0.0), _d(0.0)
    _check(getattr(lib(), "tmvx_lu_diag_range_" + suf)(n, LU.data_ptr(), lda, ctypes.byref(mn),
                                                      ctypes.byref(mx), _stream(stream)),
           "lu_diag_range")
    return mn.value, mx.value


def logdet(LU, piv, stream=None):
    """(log|det A|, sign) from the factors (TMV LUDiv::logDet); sign is
    +1, -1, or 0 for a singular matrix (logdet -inf)."""
    n = _square(LU, "LU")
    suf = _suffix(LU.dtype)
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    _on_gpu(LU, piv)
    ld, sg = _d(0.0), _d(0.0)
    fn = getattr(lib(), "tmvx_logdet_" + suf)
    _check(fn(n, LU.data_ptr(), lda, piv.data_ptr(), ctypes.byref(ld), ctypes.byref(sg),
              _stream(stream)), "logdet")
    return ld.value, sg.value


def det(LU, piv, stream=None):
    """det A from the factors, formed as TMV's LUDiv::det does:
    sign * exp(logdet), 0 when a diagonal is 0."""
    ld, sg = logdet(LU, piv, stream)
    return 0.0 if sg == 0.0 else sg * math.exp(ld)


def trsm(A, B, upper, trans=False, unit=False, stream=None):
    """B := op(A)^-1 B for triangular A (TMV TriLDivEq); raises
    TmvxError(TMVX_ERR_SINGULAR) on a zero diagonal met by a non-zero
    entry."""
    n = _square(A, "A")
    suf = _suffix(A.dtype)
    lda = _colmajor_ld(A, "A")
    Bm, ldb = _rhs(B, n, A.dtype)
    _on_gpu(A, B)
    fn = getattr(lib(), "tmvx_trsm_" + suf)
    _check(fn(int(bool(upper)), int(bool(trans)), int(bool(unit)), n, Bm.shape[1],
              A.data_ptr(), lda, Bm.data_ptr(), ldb, _stream(stream)), "trsm")
    return B


def trsm_blockinv(A, B, upper, trans=False, unit=False, stream=None):
    """B := op(A)^-1 B with the mixed solver's correction solve (64 x 64
    diagonal blocks through their inverses; not TMV's rounding).  Raises
    TmvxError(TMVX_ERR_SINGULAR) on any zero diagonal."""
    n = _square(A, "A")
    suf = _suffix(A.dtype)
    lda = _colmajor_ld(A, "A")
    Bm, ldb = _rhs(B, n, A.dtype)
    _on_gpu(A, B)
    fn = getattr(lib(), "tmvx_trsm_blockinv_" + suf)
    _check(fn(int(bool(upper)), int(bool(trans)), int(bool(unit)), n, Bm.shape[1],
              A.data_ptr(), lda, Bm.data_ptr(), ldb, _stream(stream)), "trsm_blockinv")
    return B


def laswp(A, piv, k1=0, k2=None, forward=True, stream=None):
    """Interchanges piv[k1:k2] applied to the rows of A in order (TMV
    permuteRows) or in reverse order (reversePermuteRows)."""
    m, ncols = A.shape if A.dim() == 2 else (A.shape[0], 1)
    k2 = piv.numel() if k2 is None else k2
    suf = _suffix(A.dtype)
    if not 0 <= k1 <= k2 <= m:
        raise ValueError(f"need 0 <= k1 <= k2 <= {m} (got k1={k1}, k2={k2})")
    _check_piv(piv, k2)
    if A.dim() == 2 and ncols > 1:
        lda = _colmajor_ld(A, "A")
    else:
        if m > 1 and A.stride(0) != 1:
            raise ValueError("A must have unit row stride")
        lda = max(m, 1)
    _on_gpu(A, piv)
    fn = getattr(lib(), "tmvx_laswp_" + suf)
    _check(fn(int(bool(forward)), ncols, A.data_ptr(), lda, k1, k2, piv.data_ptr(),
              _stream(stream)), "laswp")
    return A


class IrInfo(ctypes.Structure):
    """struct tmvx_ir_info (include/tmvx.h)."""
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("factor_info", ctypes.c_int), ("inner_iterations", ctypes.c_int),
                ("backward_error", ctypes.c_double), ("factor_ms", ctypes.c_double),
                ("refine_ms", ctypes.c_double), ("pivoting", ctypes.c_int),
                ("reserved", ctypes.c_int)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}
        d["pivoting"] = {PIVOT_RBT: "rbt", PIVOT_PARTIAL: "partial",
                         PIVOT_RBT | PIVOT_PARTIAL: "rbt->partial"}.get(d["pivoting"],
                                                                       d["pivoting"])
        return d


LOWPREC_F32 = 0
LOWPREC_F16 = 1
LOWPREC_BF16 = 2
LOWPREC_F16X2 = 3      # split operands (hi + lo) on FP16 MFMA
LOWPREC_BF16X2 = 4     # split operands (hi + lo) on BF16 MFMA
REFINE_GMRES = 16      # or-ed into lowprec: GMRES-IR (default for F16 / BF16 / BF16X2)
REFINE_CLASSIC = 32    # or-ed into lowprec: classical refinement (default for F32)
PIVOT_PARTIAL = 64     # or-ed into lowprec: partial pivoting (the default)
PIVOT_RBT = 128        # or-ed into lowprec: random butterfly + no interchanges,
                       # falling back to partial pivoting


def gesv_mixed(A, B, X=None, tol=1e-12, max_iter=30, lowprec=LOWPREC_F32, check=True,
               stream=None):
    """Mixed-precision solve with iterative refinement: A (FP64, column-major,
    not modified) factorised in FP32 on MFMA, residual and update in FP64.
    Returns (X, info dict).  Raises on failure when check is set."""
    torch = _torch()
    n = _square(A, "A")
    if A.dtype != torch.float64:
        raise TypeError("A must be float64")
    lda = _colmajor_ld(A, "A")
    Bm, ldb = _rhs(B, n, torch.float64)
    nrhs = Bm.shape[1]
    if X is not None:
        Xm, ldx = _rhs(X, n, torch.float64, "X")
        if Xm.shape[1] != nrhs:
            raise ValueError(f"X has {Xm.shape[1]} columns, expected {nrhs}")
    _on_gpu(A, B, *(() if X is None else (X,)))
    if X is None:
        X = empty_colmajor(n, nrhs, torch.float64, A.device)
        Xm, ldx = _rhs(X, n, torch.float64, "X")
    info = IrInfo()
    rc = lib().tmvx_gesv_mixed_f64(lowprec, n, nrhs, A.data_ptr(), lda, Bm.data_ptr(), ldb,
                                   Xm.data_ptr(), ldx, tol, max_iter, ctypes.byref(info),
                                   _stream(stream))
    if check:
        _check(rc, "gesv_mixed")
    d = info.as_dict()
    d["status"] = rc
    return (X if B.dim() > 1 else Xm.view(n)), d


def _fill_target(out) -> None:
    if out.dtype != _torch().float64 or not out.is_contiguous():
        raise TypeError("a contiguous float64 tensor is filled")
    _on_gpu(out)


def fill_uniform(out, seed, lo=-1.0, hi=1.0, offset=0, stream=None):
    """Fill a contiguous float64 device buffer (flat counter order) with the
    documented SplitMix64 uniform stream (see tmv_amd.rng)."""
    _fill_target(out)
    _check(lib().tmvx_fill_uniform_f64(seed, offset, out.numel(), lo, hi, out.data_ptr(),
                                       _stream(stream)), "fill_uniform")
    return out


def fill_normal(out, seed, offset=0, stream=None):
    _fill_target(out)
    _check(lib().tmvx_fill_normal_f64(seed, offset, out.numel(), out.data_ptr(),
                                      _stream(stream)), "fill_normal")
    return out


# ---------------------------------------------------------------------------
# Batches of small systems (include/tmvx_batched.h)
# ---------------------------------------------------------------------------

BATCH_NMAX = 128
BATCH_RHSMAX = 16


class Batch:
    """A batch of independent systems packed in one device buffer each for
    A (n_i x n_i), B (n_i x nrhs_i), the pivots and info, column-major with
    leading dimension n_i, plus the device pointer arrays the C ABI takes.
    ``a(i)`` / ``b(i)`` / ``piv(i)`` are views of system i."""

    def __init__(self, ns, nrhs=None, dtype=None, device="cuda", A=None, B=None):
        """A, B: existing flat device buffers to wrap (else allocated)."""
        torch = _torch()
        import numpy as np
        dtype = torch.float64 if dtype is None else dtype
        self.dtype = dtype
        self.ns = np.asarray(ns, dtype=np.int32)
        self.nrhs = (np.zeros_like(self.ns) if nrhs is None
                     else np.asarray(nrhs, dtype=np.int32))
        if self.nrhs.shape != self.ns.shape:
            raise ValueError("ns and nrhs differ in length")
        cnt = self.count = len(self.ns)
        na = self.ns.astype(np.int64) ** 2
        nb = self.ns.astype(np.int64) * self.nrhs
        self.offA = np.concatenate([[0], np.cumsum(na)[:-1]]).astype(np.int64) if cnt else na
        self.offB = np.concatenate([[0], np.cumsum(nb)[:-1]]).astype(np.int64) if cnt else nb
        self.offP = (np.concatenate([[0], np.cumsum(self.ns)[:-1]]).astype(np.int64)
                     if cnt else self.ns.astype(np.int64))
        self.A = (torch.zeros(max(int(na.sum()), 1), dtype=dtype, device=device)
                  if A is None else A)
        self.B = (torch.zeros(max(int(nb.sum()), 1), dtype=dtype, device=device)
                  if B is None else B)
        if self.A.numel() < int(na.sum()) or self.B.numel() < int(nb.sum()):
            raise ValueError("buffer smaller than the batch layout")
        self.P = torch.zeros(max(int(self.ns.sum()), 1), dtype=torch.int64, device=device)
        self.info = torch.zeros(max(cnt, 1), dtype=torch.int32, device=device)
        es = self.A.element_size()
        self.d_n = torch.as_tensor(self.ns, device=device)
        self.d_nrhs = torch.as_tensor(self.nrhs, device=device)
        self.pA = torch.as_tensor(self.A.data_ptr() + es * self.offA, device=device)
        self.pB = torch.as_tensor(self.B.data_ptr() + es * self.offB, device=device)
        self.pP = torch.as_tensor(self.P.data_ptr() + 8 * self.offP, device=device)

    def a(self, i):
        n = int(self.ns[i]); o = int(self.offA[i])
        return self.A[o:o + n * n].view(n, n).t()

    def b(self, i):
        n, r = int(self.ns[i]), int(self.nrhs[i]); o = int(self.offB[i])
        return self.B[o:o + n * r].view(r, n).t()

    def piv(self, i):
        n = int(self.ns[i]); o = int(self.offP[i])
        return self.P[o:o + n]


class BatchPlan:
    """tmvx_batch_plan: the work list and schedule of one batch shape."""

    def __init__(self, batch_or_ns, nrhs=None, dtype=None, tune=False, stream=None):
        torch = _torch()
        import numpy as np
        if isinstance(batch_or_ns, Batch):
            ns, nrhs, dtype = batch_or_ns.ns, batch_or_ns.nrhs, batch_or_ns.dtype
        else:
            ns = np.asarray(batch_or_ns, dtype=np.int32)
            nrhs = np.zeros_like(ns) if nrhs is None else np.asarray(nrhs, dtype=np.int32)
        dtype = torch.float64 if dtype is None else dtype
        self.dtype = dtype
        self._ns = np.ascontiguousarray(ns, dtype=np.int32)
        self._nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
        h = _vp()
        rc = lib().tmvx_batch_plan_create(0 if dtype == torch.float64 else 1, len(self._ns),
                                          self._ns.ctypes.data, self._nrhs.ctypes.data,
                                          int(bool(tune)), ctypes.byref(h), _stream(stream))
        _check(rc, "batch_plan_create")
        self.handle = h

    def describe(self) -> str:
        return lib().tmvx_batch_plan_describe(self.handle).decode()

    def set(self, schedule, split, wgs_per_cu, stream=None):
        _check(lib().tmvx_batch_plan_set(self.handle, schedule, split, wgs_per_cu,
                                         _stream(stream)), "batch_plan_set")

    def close(self):
        if getattr(self, "handle", None):
            lib().tmvx_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_schedule_count() -> int:
    return lib().tmvx_batch_schedule_count()


def gesv_batched(plan: BatchPlan, batch: Batch, stream=None):
    """Factor every A_i in place and solve A_i X_i = B_i (B overwritten), in
    one launch.  Returns the per-system info tensor (device)."""
    fn = getattr(lib(), "tmvx_gesv_batched_" + _suffix(batch.dtype))
    _check(fn(plan.handle, batch.pA.data_ptr(), batch.pP.data_ptr(), batch.pB.data_ptr(),
              batch.info.data_ptr(), _stream(stream)), "gesv_batched")
    return batch.info


def getrs_batched(batch: Batch, trans=False, stream=None):
    """Solve with the stored factors of ``batch`` (B overwritten)."""
    fn = getattr(lib(), "tmvx_getrs_batched_" + _suffix(batch.dtype))
    _check(fn(int(bool(trans)), batch.count, batch.d_n.data_ptr(), batch.d_nrhs.data_ptr(),
              batch.pA.data_ptr(), batch.pP.data_ptr(), batch.pB.data_ptr(),
              _stream(stream)), "getrs_batched")


def gemm_batched(m, n, k, pA, pB, pC, alpha=1.0, beta=0.0, dtype=None, stream=None):
    """C_i = alpha A_i B_i + beta C_i over device size arrays m, n, k (int32)
    and device pointer arrays pA, pB, pC (int64)."""
    torch = _torch()
    dtype = torch.float64 if dtype is None else dtype
    fn = getattr(lib(), "tmvx_gemm_batched_" + _suffix(dtype))
    _check(fn(m.numel(), m.data_ptr(), n.data_ptr(), k.data_ptr(), alpha, pA.data_ptr(),
              pB.data_ptr(), beta, pC.data_ptr(), _stream(stream)), "gemm_batched")


def getrf_batched(batch: Batch, stream=None):
    """Factor every system of ``batch`` in place (any n: n <= 128 batched,
    larger ones through the blocked LU).  Synchronises; returns info."""
    import numpy as np
    fn = getattr(lib(), "tmvx_getrf_batched_" + _suffix(batch.dtype))
    nh = np.ascontiguousarray(batch.ns, dtype=np.int32)
    _check(fn(batch.count, nh.ctypes.data, batch.pA.data_ptr(), batch.pP.data_ptr(),
              batch.info.data_ptr(), _stream(stream)), "getrf_batched")
    return batch.info


def gesv_adjoint(LU, piv, X, G, gradA=None, accumulate=False, stream=None):
    """gradB = A^-T G, gradA (+)= -gradB X^T for one system with the factors
    of getrf.  X, G n x nrhs column-major (1-D = one column).  Returns
    (gradB, gradA)."""
    n = _square(LU, "LU")
    suf = _suffix(LU.dtype)
    lda = _colmajor_ld(LU, "LU")
    _check_piv(piv, n)
    Xm, ldx = _rhs(X, n, LU.dtype, "X")
    Gm, ldg = _rhs(G, n, LU.dtype, "G")
    nrhs = Xm.shape[1]
    if Gm.shape[1] != nrhs:
        raise ValueError(f"G has {Gm.shape[1]} columns, X has {nrhs}")
    if gradA is not None:
        if gradA.dtype != LU.dtype or tuple(gradA.shape) != (n, n):
            raise ValueError(f"gradA must be a {n} x {n} {LU.dtype} tensor")
    _on_gpu(LU, piv, X, G, *(() if gradA is None else (gradA,)))
    gB = empty_colmajor(n, nrhs, LU.dtype, LU.device)
    if gradA is None:
        gradA = empty_colmajor(n, n, LU.dtype, LU.device)
        accumulate = False
    ldga = _colmajor_ld(gradA, "gradA")
    fn = getattr(lib(), "tmvx_gesv_adjoint_" + suf)
    _check(fn(n, nrhs, LU.data_ptr(), lda, piv.data_ptr(), Xm.data_ptr(), ldx, Gm.data_ptr(),
              ldg, gB.data_ptr(), max(n, 1), gradA.data_ptr(), ldga, int(bool(accumulate)),
              _stream(stream)), "gesv_adjoint")
    return (gB if X.dim() > 1 else gB.view(n)), gradA


def gesv_adjoint_batched(fact: Batch, X: Batch, G: Batch, gradB: Batch, gradA: Batch,
                         accumulate=False, recompute=False, stream=None):
    """Batched adjoint.  ``fact`` holds the factors in fact.A / fact.P
    (recompute=False) or the original matrices in fact.A (recompute=True);
    X, G, gradB are batches whose B buffers hold the n_i x nrhs_i blocks;
    gradA's A buffer receives the n_i x n_i gradients."""
    import numpy as np
    suf = _suffix(fact.dtype)
    if recompute:
        fn = getattr(lib(), "tmvx_gesv_adjoint_batched_recompute_" + suf)
        nh = np.ascontiguousarray(fact.ns, dtype=np.int32)
        rh = np.ascontiguousarray(X.nrhs, dtype=np.int32)
        _check(fn(fact.count, nh.ctypes.data, rh.ctypes.data, fact.pA.data_ptr(),
                  X.pB.data_ptr(), G.pB.data_ptr(), gradB.pB.data_ptr(), gradA.pA.data_ptr(),
                  int(bool(accumulate)), _stream(stream)), "gesv_adjoint_batched_recompute")
    else:
        fn = getattr(lib(), "tmvx_gesv_adjoint_batched_" + suf)
        _check(fn(fact.count, fact.d_n.data_ptr(), X.d_nrhs.data_ptr(), fact.pA.data_ptr(),
                  fact.pP.data_ptr(), X.pB.data_ptr(), G.pB.data_ptr(), gradB.pB.data_ptr(),
                  gradA.pA.data_ptr(), int(bool(accumulate)), _stream(stream)),
               "gesv_adjoint_batched")


def gemm_lowp(A, B, C=None, alpha=1.0, beta=0.0, lowprec=LOWPREC_F16, scale_b=True, stream=None):
    """C := alpha A B + beta C for column-major float32 tensors on FP16 / BF16
    MFMA with FP32 accumulation (tmvx_gemm_f32_lowp)."""
    torch = _torch()
    m, k = A.shape
    k2, n = B.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    if A.dtype != torch.float32 or B.dtype != torch.float32 or (
            C is not None and C.dtype != torch.float32):
        raise TypeError("float32 operands only")
    if C is not None and tuple(C.shape) != (m, n):
        raise ValueError(f"C has shape {tuple(C.shape)}, expected {(m, n)}")
    lda = _colmajor_ld(A, "A")
    ldb = _colmajor_ld(B, "B")
    _on_gpu(A, B, *(() if C is None else (C,)))
    if C is None:
        C = empty_colmajor(m, n, torch.float32, A.device)
        beta = 0.0
    ldc = _colmajor_ld(C, "C")
    _check(lib().tmvx_gemm_f32_lowp(lowprec, m, n, k, alpha, A.data_ptr(), lda, B.data_ptr(),
                                    ldb, beta, C.data_ptr(), ldc, int(bool(scale_b)),
                                    _stream(stream)), "gemm_lowp")
    return C
