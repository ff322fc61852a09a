"""The oracle: CPU restatement of TMV 0.76's dense GEMM + LU path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of bench.py import this package, and only as the checker.
The product library (``tmv_amd``) never imports it and has no CPU fallback.

Two native libraries back it:

* ``oracle/liboracle.so`` -- ``tmv_oracle.cpp``, a fresh restatement of
  TMV's algorithms (MultMM Block path, RecursiveLUDecompose, LU_LDivEq,
  NonBlasTriLDivEq), built by :func:`build`.
* ``oracle/_ref/libtmv_ref.so`` -- the real TMV library compiled from the
  reference sources by ``oracle/Makefile.ref`` plus ``tmvref_driver.cpp``.
  Present only where it was built (it travels with the tree as a binary).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtmv_ref.so")
REF_SRC = "/root/reference"

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p


def build(verbose: bool = False) -> None:
    """Compile the restatement (always) and the real TMV (when its sources
    are present in this container)."""
    cmd = ["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++14",
           os.path.join(HERE, "tmv_oracle.cpp"), "-o", ORACLE_SO + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(ORACLE_SO + ".tmp", ORACLE_SO)
    if os.path.isdir(os.path.join(REF_SRC, "src")):
        root = os.path.dirname(HERE)
        subprocess.run(["make", "-f", os.path.join("oracle", "Makefile.ref"),
                        "REF=" + REF_SRC, "-j8"], cwd=root, check=True,
                       capture_output=not verbose)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class _Lib:
    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        self.path = path

    def fn(self, name: str):
        return getattr(self.lib, self.prefix + name)


_oracle = None
_ref = None


def oracle_lib() -> _Lib:
    global _oracle
    if _oracle is None:
        _oracle = _Lib(ORACLE_SO, "orc_")
    return _oracle


def ref_lib() -> _Lib | None:
    """The real TMV build, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = _Lib(REF_SO, "tmvref_")
    return _ref


_SUF = {np.dtype(np.float64): ("f64", ctypes.c_double),
        np.dtype(np.float32): ("f32", ctypes.c_float),
        np.dtype(np.longdouble): ("f80", ctypes.c_longdouble)}


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _fcopy(a, dtype):
    return np.array(a, dtype=dtype, order="F", copy=True)


def gemm(A, B, C=None, alpha=1.0, add=False, a_rowmajor=False,
         b_rowmajor=False, dtype=np.float64, lib: _Lib | None = None):
    """C (+)= alpha*A*B following tmv::MultMM<add>.  A and B are taken in the
    stated storage order (row-major = C order), C is column-major."""
    lib = lib or oracle_lib()
    dtype = np.dtype(dtype)
    suf, ct = _SUF[dtype]
    A = np.ascontiguousarray(A, dtype) if a_rowmajor else _fortran(A, dtype)
    B = np.ascontiguousarray(B, dtype) if b_rowmajor else _fortran(B, dtype)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), dtype, order="F") if C is None else _fcopy(C, dtype)
    lda = A.shape[1] if a_rowmajor else max(1, m)
    ldb = B.shape[1] if b_rowmajor else max(1, k)
    f = lib.fn("gemm_" + suf)
    f.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _i64, _i64, ct, _vp, _i64,
                  _vp, _i64, ctypes.c_int, _vp, _i64]
    rc = f(int(a_rowmajor), int(b_rowmajor), m, n, k, alpha, _ptr(A), lda,
           _ptr(B), ldb, int(add), _ptr(C), max(1, m))
    if rc != 0:
        raise RuntimeError(f"gemm failed: {rc}")
    return C


def gemm_complex(A, B, C=None, alpha=1.0, add=False, conj_a=False, conj_b=False,
                 a_rowmajor=False, b_rowmajor=False, dtype=np.complex128):
    """Complex C (+)= alpha*op(A)*op(B) as TMV computes it: four real products
    of the real and imaginary parts, conjugations folded into the signs
    (DoBlockMultMM's complex overloads, src/TMV_MultMM_Block.cpp:370-392),
    each real product by the real restatement above.  A complex alpha scales
    each real product into both parts of C (the real-product, complex-C
    kernel adds x * tile)."""
    rdt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    A = np.asarray(A)
    B = np.asarray(B)
    Ar, Ai = np.ascontiguousarray(A.real), np.ascontiguousarray(A.imag)
    Br, Bi = np.ascontiguousarray(B.real), np.ascontiguousarray(B.imag)
    m, n = A.shape[0], B.shape[1]
    if C is None:
        Cr = np.zeros((m, n), rdt)
        Ci = np.zeros((m, n), rdt)
    else:
        Cr, Ci = np.asarray(C).real.copy(), np.asarray(C).imag.copy()
    kw = dict(a_rowmajor=a_rowmajor, b_rowmajor=b_rowmajor, dtype=rdt)
    x = complex(alpha)
    if x.imag == 0.0:
        xr = x.real
        Cr = gemm(Ar, Br, C=Cr, alpha=xr, add=add, **kw)
        Cr = gemm(Ai, Bi, C=Cr, alpha=-xr if conj_a == conj_b else xr, add=True, **kw)
        Ci = gemm(Ar, Bi, C=Ci, alpha=-xr if conj_b else xr, add=add, **kw)
        Ci = gemm(Ai, Br, C=Ci, alpha=-xr if conj_a else xr, add=True, **kw)
    else:
        ix = 1j * x
        terms = [(Ar, Br, x, add), (Ai, Bi, -x if conj_a == conj_b else x, True),
                 (Ar, Bi, -ix if conj_b else ix, True), (Ai, Br, -ix if conj_a else ix, True)]
        for P, Q, y, ad in terms:
            Cr = gemm(P, Q, C=Cr, alpha=y.real, add=ad, **kw)
            Ci = gemm(P, Q, C=Ci, alpha=y.imag, add=ad, **kw)
    out = np.empty((m, n), np.dtype(dtype), order="F")
    out.real, out.imag = Cr, Ci
    return out


def ref_gemm_complex(A, B, C=None, alpha=1.0, add=False, conj_a=False, conj_b=False,
                     a_rowmajor=False, b_rowmajor=False):
    """Complex MultMM<add> through the real TMV (complex<double>)."""
    lib = ref_lib()
    if lib is None:
        raise FileNotFoundError(REF_SO)
    A = np.ascontiguousarray(A, np.complex128) if a_rowmajor else np.asfortranarray(A, np.complex128)
    B = np.ascontiguousarray(B, np.complex128) if b_rowmajor else np.asfortranarray(B, np.complex128)
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros((m, n), np.complex128, order="F") if C is None else np.array(C, np.complex128, order="F")
    lda = A.shape[1] if a_rowmajor else max(1, m)
    ldb = B.shape[1] if b_rowmajor else max(1, k)
    x = complex(alpha)
    f = lib.fn("gemm_z")
    f.argtypes = [ctypes.c_int] * 4 + [_i64] * 3 + [ctypes.c_double] * 2 + \
        [_vp, _i64, _vp, _i64, ctypes.c_int, _vp, _i64]
    rc = f(int(a_rowmajor), int(b_rowmajor), int(conj_a), int(conj_b), m, n, k, x.real, x.imag,
           _ptr(A), lda, _ptr(B), ldb, int(add), _ptr(C), max(1, m))
    if rc != 0:
        raise RuntimeError(f"complex gemm failed: {rc}")
    return C


def getrf(A, dtype=np.float64, gaps=False, lib: _Lib | None = None):
    """In-place PA = LU (TMV LU_Decompose).  Returns (LU, P) with P the
    0-based interchange list; with gaps=True also the per-column relative
    pivot gap (oracle restatement only)."""
    lib = lib or oracle_lib()
    dtype = np.dtype(dtype)
    suf, _ = _SUF[dtype]
    LU = _fcopy(A, dtype)
    m, n = LU.shape
    P = np.zeros(min(m, n), np.int64)
    g = np.zeros(min(m, n), np.float64)
    f = lib.fn("getrf_" + suf)
    if lib.prefix == "orc_":
        f.argtypes = [_i64, _i64, _vp, _i64, _vp, _vp]
        rc = f(m, n, _ptr(LU), max(1, m), _ptr(P), _ptr(g) if gaps else None)
    elif m != n:
        f = lib.fn("getrf_mn_" + suf)
        f.argtypes = [_i64, _i64, _vp, _i64, _vp]
        P = np.zeros(max(m, n), np.int64)     # TMV writes at most max(m, n) entries
        rc = f(m, n, _ptr(LU), max(1, m), _ptr(P))
        P = P[:min(m, n)].copy()
    else:
        f.argtypes = [_i64, _vp, _i64, _vp]
        rc = f(n, _ptr(LU), max(1, n), _ptr(P))
    if rc != 0:
        raise RuntimeError(f"getrf failed: {rc}")
    return (LU, P, g) if gaps else (LU, P)


class SingularError(RuntimeError):
    """A zero diagonal met a non-zero right-hand side (TMV's Singular)."""


def getrf_complex(A, gaps=False, lib: _Lib | None = None):
    """Complex (complex128) PA = LU: the oracle's restatement, or the real
    TMV with lib=ref_lib().  Returns (LU, P[, gaps])."""
    lib = lib or oracle_lib()
    A = np.array(A, np.complex128, order="F")
    n = A.shape[0]
    P = np.zeros(n, np.int64)
    g = np.zeros(n, np.float64)
    if lib is oracle_lib():
        f = lib.fn("getrf_z")
        f.argtypes = [_i64, _i64, _vp, _i64, _vp, _vp]
        rc = f(n, n, _ptr(A), max(1, n), _ptr(P), _ptr(g) if gaps else None)
    else:
        f = lib.fn("getrf_z")
        f.argtypes = [_i64, _vp, _i64, _vp]
        rc = f(n, _ptr(A), max(1, n), _ptr(P))
    if rc != 0:
        raise RuntimeError(f"complex getrf failed: {rc}")
    return (A, P, g) if gaps else (A, P)


def getrs_complex(LU, P, B, trans=False):
    """Complex solve with the factors (LU_LDivEq; trans: LU_RDivEq)."""
    lib = oracle_lib()
    LU = np.asfortranarray(LU, np.complex128)
    n = LU.shape[0]
    B = np.asarray(B, np.complex128)
    vec = B.ndim == 1
    X = np.array(B.reshape(n, -1), np.complex128, order="F")
    P = np.ascontiguousarray(P, np.int64)
    f = lib.fn("getrs_z")
    f.argtypes = [ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _vp, _i64]
    rc = f(int(trans), n, X.shape[1], _ptr(LU), max(1, n), _ptr(P), _ptr(X), max(1, n))
    if rc == -2:
        raise SingularError("singular upper triangle")
    if rc != 0:
        raise RuntimeError(f"complex getrs failed: {rc}")
    return X[:, 0] if vec else X


def householder_reflect(x):
    """TMV's HouseholderReflect (src/TMV_Householder.cpp:46-180) on a real
    vector x (x[0] = x0), in place: returns beta; x[0] becomes y*scale and
    x[1:] the reflector (its implicit first element 1)."""
    rmin = np.finfo(x.dtype).tiny
    tail = x[1:]
    scale = max(float(np.abs(tail).max(initial=0.0)), abs(float(x[0])))
    if scale < rmin:                       # hopeless: zero the vector
        x[:] = 0.0
        return x.dtype.type(0)
    invscale = 1.0 / scale
    normsqx = float(np.sum((tail * invscale) ** 2))
    if normsqx == 0.0:                     # H = I
        tail[:] = 0.0
        return x.dtype.type(0)
    x0 = float(x[0]) * invscale
    normsq = normsqx + x0 * x0
    normx = np.sqrt(normsq)
    y = -normx if x0 > 0 else normx
    v0 = x0 - y
    beta = (v0 * v0) / (normsq - y * x0)
    tail *= (1.0 / v0) * invscale
    x[0] = y * scale
    return x.dtype.type(beta)


def qr_decompose(A, dtype=np.float64):
    """TMV's NonBlockQRDecompose (src/TMV_QRDecompose.cpp:55-110): per
    column the reflector (householder_reflect), then H = I - beta v v^T on
    the columns to its right (HouseholderLMult, src/TMV_Householder.cpp:
    333-390).  Returns (QR packed like TMV's, beta)."""
    A = np.array(A, dtype, order="F")
    m, n = A.shape
    beta = np.zeros(n, dtype)
    for j in range(n):
        b = householder_reflect(A[j:, j])
        beta[j] = b
        if b != 0 and j + 1 < n:
            v = np.concatenate(([1.0], A[j + 1:, j])).astype(dtype)
            w = v @ A[j:, j + 1:]
            A[j:, j + 1:] -= b * np.outer(v, w)
    return A, beta


def qr_solve(QR, beta, B):
    """QR_LDiv (src/TMV_QRDiv.cpp:33-94): y = Q^T b (the reflectors in
    order), then x = R^-1 y(0:n)."""
    QR = np.asarray(QR)
    m, n = QR.shape
    Y = np.array(B, np.float64).reshape(m, -1)
    for j in range(n):
        if beta[j] != 0:
            v = np.concatenate(([1.0], QR[j + 1:, j]))
            Y[j:] -= beta[j] * np.outer(v, v @ Y[j:])
    R = np.triu(QR[:n, :n])
    if np.any(np.diag(R) == 0):
        raise SingularError("singular R")
    X = sla.solve_triangular(R, Y[:n], lower=False)
    return X[:, 0] if np.asarray(B).ndim == 1 else X


def ref_geqrf(A):
    """QR_Decompose through the real TMV: (QR, beta)."""
    lib = ref_lib()
    if lib is None:
        raise FileNotFoundError(REF_SO)
    A = np.array(A, np.float64, order="F")
    m, n = A.shape
    beta = np.zeros(n)
    f = lib.fn("geqrf_f64")
    f.argtypes = [_i64, _i64, _vp, _i64, _vp]
    if f(m, n, _ptr(A), max(1, m), _ptr(beta)) != 0:
        raise RuntimeError("geqrf failed")
    return A, beta


def ref_qr_solve(A, B):
    """x = b/A with QR division through the real TMV."""
    lib = ref_lib()
    if lib is None:
        raise FileNotFoundError(REF_SO)
    A = np.asfortranarray(A, np.float64)
    m, n = A.shape
    Bv = np.asarray(B, np.float64)
    Bm = np.asfortranarray(Bv.reshape(m, -1))
    X = np.zeros((n, Bm.shape[1]), order="F")
    f = lib.fn("qr_solve_f64")
    f.argtypes = [_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]
    rc = f(m, n, Bm.shape[1], _ptr(A), max(1, m), _ptr(Bm), max(1, m), _ptr(X), max(1, n))
    if rc == -2:
        raise SingularError("singular")
    if rc != 0:
        raise RuntimeError(f"qr solve failed: {rc}")
    return X[:, 0] if Bv.ndim == 1 else X


def getrs(LU, P, B, trans=False, dtype=np.float64, lib: _Lib | None = None):
    """Solve with the factors (TMV LU_LDivEq; trans -> LU_RDivEq)."""
    lib = lib or oracle_lib()
    dtype = np.dtype(dtype)
    suf, _ = _SUF[dtype]
    LU = _fortran(LU, dtype)
    n = LU.shape[0]
    B = np.asarray(B, dtype)
    vec = B.ndim == 1
    X = _fcopy(B.reshape(n, -1), dtype)
    P = np.ascontiguousarray(P, np.int64)
    f = lib.fn("getrs_" + suf)
    f.argtypes = [ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _vp, _i64]
    rc = f(int(trans), n, X.shape[1], _ptr(LU), max(1, n), _ptr(P), _ptr(X),
           max(1, n))
    if rc == -2:
        raise SingularError("singular upper triangle")
    if rc != 0:
        raise RuntimeError(f"getrs failed: {rc}")
    return X[:, 0] if vec else X


def gesv(A, B, dtype=np.float64):
    """x = b/A the TMV way: factor a copy, then solve."""
    LU, P = getrf(A, dtype)
    return getrs(LU, P, B, dtype=dtype)


def ref_gesv(A, B, dtype=np.float64):
    """x = b/A through the real TMV user API (divideUsing(LU))."""
    lib = ref_lib()
    if lib is None:
        raise FileNotFoundError(REF_SO)
    dtype = np.dtype(dtype)
    suf, _ = _SUF[dtype]
    A = _fortran(A, dtype)
    n = A.shape[0]
    Bv = np.asarray(B, dtype)
    vec = Bv.ndim == 1
    Bm = _fortran(Bv.reshape(n, -1), dtype)
    X = np.zeros_like(Bm, order="F")
    f = lib.fn("gesv_" + suf)
    f.argtypes = [_i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]
    rc = f(n, Bm.shape[1], _ptr(A), n, _ptr(Bm), n, _ptr(X), n)
    if rc == -2:
        raise SingularError("singular")
    if rc != 0:
        raise RuntimeError(f"gesv failed: {rc}")
    return X[:, 0] if vec else X


def ref_det(A):
    lib = ref_lib()
    A = _fortran(A, np.float64)
    out = ctypes.c_double(0.0)
    f = lib.fn("det_f64")
    f.argtypes = [_i64, _vp, _i64, ctypes.c_void_p]
    rc = f(A.shape[0], _ptr(A), A.shape[0], ctypes.addressof(out))
    if rc != 0:
        raise RuntimeError("det failed")
    return out.value


def det_from_lu(LU, P) -> float:
    f = oracle_lib().fn("det_from_lu_f64")
    f.restype = ctypes.c_double
    f.argtypes = [_i64, _vp, _i64, _vp]
    LU = _fortran(LU, np.float64)
    P = np.ascontiguousarray(P, np.int64)
    return f(LU.shape[0], _ptr(LU), LU.shape[0], _ptr(P))


def logdet_from_lu(LU, P) -> tuple[float, float]:
    """(log|det|, sign) from the factors (LUDiv::logDet)."""
    f = oracle_lib().fn("logdet_from_lu_f64")
    f.restype = ctypes.c_double
    f.argtypes = [_i64, _vp, _i64, _vp, ctypes.c_void_p]
    LU = _fortran(LU, np.float64)
    P = np.ascontiguousarray(P, np.int64)
    sign = ctypes.c_double(0.0)
    ld = f(LU.shape[0], _ptr(LU), LU.shape[0], _ptr(P), ctypes.addressof(sign))
    return ld, sign.value


def inverse_from_lu(LU, P, dtype=np.float64):
    """A^-1 from the factors: the solve of A X = I (the matrix TMV's
    LU_Inverse returns, src/TMV_LUInverse.cpp:235-275)."""
    n = np.asarray(LU).shape[0]
    return getrs(LU, P, np.eye(n), dtype=dtype)


def ref_inverse(A):
    """A.inverse() through the real TMV (LU division)."""
    lib = ref_lib()
    if lib is None:
        raise FileNotFoundError(REF_SO)
    A = _fortran(A, np.float64)
    n = A.shape[0]
    X = np.zeros((n, n), order="F")
    f = lib.fn("inverse_f64")
    f.argtypes = [_i64, _vp, _i64, _vp, _i64]
    rc = f(n, _ptr(A), n, _ptr(X), n)
    if rc == -2:
        raise SingularError("singular")
    if rc != 0:
        raise RuntimeError(f"inverse failed: {rc}")
    return X


def ref_logdet(A) -> tuple[float, float]:
    """A.logDet(&sign) through the real TMV (LU division)."""
    lib = ref_lib()
    if lib is None:
        raise FileNotFoundError(REF_SO)
    A = _fortran(A, np.float64)
    out, sign = ctypes.c_double(0.0), ctypes.c_double(0.0)
    f = lib.fn("logdet_f64")
    f.argtypes = [_i64, _vp, _i64, ctypes.c_void_p, ctypes.c_void_p]
    if f(A.shape[0], _ptr(A), A.shape[0], ctypes.addressof(out), ctypes.addressof(sign)) != 0:
        raise RuntimeError("logdet failed")
    return out.value, sign.value


def set_threads(n: int) -> None:
    for lib in (oracle_lib(), ref_lib()):
        if lib is not None:
            lib.fn("set_threads")(ctypes.c_int(n))


def unpack_lu(LU):
    """(L unit-lower, U upper) from the packed factors."""
    LU = np.asarray(LU)
    n = min(LU.shape)
    L = np.tril(LU[:, :n], -1) + np.eye(LU.shape[0], n, dtype=LU.dtype)
    U = np.triu(LU[:n, :])
    return L, U


def trsm(A, B, upper, trans=False, unit=False, dtype=np.float64, lib: _Lib | None = None):
    """op(A)^-1 B for triangular A (TMV TriLDivEq / NonBlasTriLDivEq,
    src/TMV_TriDiv_M.cpp:190-266): the recursive split restated in
    orc_trsm_*; trans solves with A^T (the transpose is formed here, so the
    lower and upper roles swap).  Raises SingularError on a zero diagonal
    met by a non-zero entry."""
    lib = lib or oracle_lib()
    dtype = np.dtype(dtype)
    suf, _ = _SUF[dtype]
    A = np.asarray(A, dtype)
    if trans:
        A, upper = A.T, not upper
    A = _fortran(A, dtype)
    n = A.shape[0]
    B = np.asarray(B, dtype)
    vec = B.ndim == 1
    X = _fcopy(B.reshape(n, -1), dtype)
    f = lib.fn("trsm_" + suf)
    f.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _i64]
    rc = f(int(bool(upper)), int(bool(unit)), n, X.shape[1], _ptr(A), max(1, n), _ptr(X),
           max(1, n))
    if rc == -2:
        raise SingularError("singular triangle")
    if rc != 0:
        raise RuntimeError(f"trsm failed: {rc}")
    return X[:, 0] if vec else X


def apply_perm_rows(A, P):
    """Apply the interchange list in order (TMV permuteRows)."""
    A = np.array(A, copy=True)
    for i, p in enumerate(P):
        if p != i:
            A[[i, p]] = A[[p, i]]
    return A
