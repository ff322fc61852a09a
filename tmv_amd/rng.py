"""Seeded synthetic inputs (TMV itself has no RNG).

Counter-based SplitMix64: ``key = mix(seed)``, word ``i`` = ``mix(key + i)``
where ``mix`` is one SplitMix64 step (add the golden gamma 0x9E3779B97F4A7C15,
then the variant-13 finaliser).  ``u = (word >> 11) * 2**-53`` in [0, 1).
The GPU generator ``tmvx_fill_uniform_f64`` produces bit-identical uniform
values; normals (Box-Muller on word pairs 2i, 2i+1) agree to rounding.

Matrices are filled in column-major order: element (i, j) of an m x n
matrix uses counter ``offset + i + j*m``.
"""
from __future__ import annotations

import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)
_G = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def _mix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + _G
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
    return z ^ (z >> np.uint64(31))


def key(seed: int) -> np.uint64:
    return _mix(np.array([seed], dtype=np.uint64))[0]


def words(seed: int, offset: int, count: int) -> np.ndarray:
    k = key(seed)
    idx = np.arange(count, dtype=np.uint64) + np.uint64(offset)
    with np.errstate(over="ignore"):
        return _mix(idx + k)


def uniform(seed: int, count: int, lo: float = 0.0, hi: float = 1.0,
            offset: int = 0) -> np.ndarray:
    u = (words(seed, offset, count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return lo + (hi - lo) * u


def normal(seed: int, count: int, offset: int = 0) -> np.ndarray:
    w = words(seed, 2 * offset, 2 * count)
    u1 = ((w[0::2] >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (w[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def uniform_matrix(seed: int, m: int, n: int, lo: float = -1.0,
                   hi: float = 1.0) -> np.ndarray:
    """m x n column-major (Fortran-ordered) matrix of uniform entries."""
    return np.asfortranarray(uniform(seed, m * n, lo, hi).reshape(n, m).T)


def normal_matrix(seed: int, m: int, n: int) -> np.ndarray:
    return np.asfortranarray(normal(seed, m * n).reshape(n, m).T)
