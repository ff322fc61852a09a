"""Numerical checks shared by the tests (plain numpy)."""
from __future__ import annotations

import numpy as np

EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def gemm_bound_ok(C, Cref, A, B, k, eps, factor=1.0, C0=None, beta=0.0, alpha=1.0):
    """Elementwise |C - Cref| <= factor * k * eps * (|alpha| |A||B| + |beta||C0|)
    (+ a tiny absolute floor for exact zeros).  factor = 1 is the stated
    k eps |A||B| bound against an exact (long-double) product; two
    working-precision results each within it differ by at most factor 2."""
    mag = abs(alpha) * (np.abs(A) @ np.abs(B))
    if C0 is not None and beta != 0:
        mag = mag + abs(beta) * np.abs(C0)
    bound = factor * max(k, 1) * eps * mag + 1e-300
    err = np.abs(np.asarray(C, np.float64) - np.asarray(Cref, np.float64))
    return bool(np.all(err <= bound)), float((err / bound).max())


def lu_residual(A, LU, P):
    """||P A - L U||_max / (n * eps * ||A||_max) using the interchange list,
    evaluated in long double."""
    A = np.asarray(A, np.longdouble)
    LU = np.asarray(LU, np.longdouble)
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n, dtype=np.longdouble)
    U = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(np.asarray(P)):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    R = PA - L @ U
    return float(np.abs(R).max() / max(np.abs(A).max(), 1e-300))


def lu_residual_mn(A, LU, P):
    """||P A - L U||_max / ||A||_max for an m x n factorisation (L m x r
    unit-lower, U r x n upper, r = min(m, n)), in long double."""
    A = np.asarray(A, np.longdouble)
    LU = np.asarray(LU, np.longdouble)
    m, n = A.shape
    r = min(m, n)
    L = np.tril(LU[:, :r], -1) + np.eye(m, r, dtype=np.longdouble)
    U = np.triu(LU[:r, :])
    PA = A.copy()
    for i, p in enumerate(np.asarray(P)):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    R = PA - L @ U
    return float(np.abs(R).max() / max(np.abs(A).max(), 1e-300))


def backward_error(A, x, b):
    """||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), in long double."""
    A = np.asarray(A, np.longdouble)
    x = np.asarray(x, np.longdouble)
    b = np.asarray(b, np.longdouble)
    r = b - A @ x
    nA = np.abs(A).sum(axis=1).max()
    return float(np.abs(r).max() / (nA * np.abs(x).max() + np.abs(b).max()))


def pivot_gap_ok(gaps, n, eps, scale=64.0):
    """True where the pivot choice is unambiguous (gap above scale*n*eps)."""
    return np.asarray(gaps) > scale * n * eps


def inverse_residual(A, X):
    """||A X - I||_max / (n ||A||_max ||X||_max): below a few eps for an
    inverse formed by backward-stable solves (float64 product; long double
    up to n = 200)."""
    A = np.asarray(A)
    X = np.asarray(X)
    n = A.shape[0]
    if n == 0:
        return 0.0
    if n <= 200:
        R = A.astype(np.longdouble) @ X.astype(np.longdouble) - np.eye(n, dtype=np.longdouble)
    else:
        R = A.astype(np.float64) @ X.astype(np.float64) - np.eye(n)
    den = n * max(float(np.abs(A).max()), 1e-300) * max(float(np.abs(X).max()), 1e-300)
    return float(np.abs(R).max()) / den


def inverse_close(X, Xref, A, eps, factor=64.0):
    """||X - Xref||_max <= factor n eps kappa ||Xref||_max with kappa =
    ||A||_1 ||Xref||_1: two inverses from backward-stable LU agree to the
    conditioning of A."""
    n = A.shape[0]
    kappa = np.abs(A).sum(axis=0).max() * np.abs(Xref).sum(axis=0).max()
    bound = factor * max(n, 1) * eps * kappa * np.abs(Xref).max()
    err = float(np.abs(np.asarray(X, np.float64) - Xref).max())
    return err <= bound, err / bound
