"""Known-answer fixtures held by TMV's own tests and examples, as data.

Each builder returns the matrix (float64, numpy) exactly as the cited TMV
test or example constructs it, and KNOWN holds the answers those files
state.  Citations are to the TMV source tree.
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(HERE, "known_answers.json")) as f:
    KNOWN = json.load(f)


def det_matrices() -> dict[str, np.ndarray]:
    """tests/TMV_TestMatrixDet.cpp:9-84 (the `<<` lists are row-major)."""
    out = {}
    for name, rows in KNOWN["det_matrices"].items():
        out[name] = np.array(rows, dtype=np.float64)
    m5 = out["m5"]
    out["m9"] = np.outer(m5[0], m5[0])                  # m5.row(0) ^ m5.row(0)
    out["m10"] = m5[:4].T @ m5[:4]                      # rowRange(0,4)^T * ...
    return out


def square_div_4x4() -> tuple[np.ndarray, np.ndarray]:
    """tests/TMV_TestMatrixDiv.cpp:11-24: m(i,j) = 2+4i-5j with five pinned
    entries, b = (2,-10,5,-5); det(m) = 28800 (:76)."""
    m = np.fromfunction(lambda i, j: 2 + 4 * i - 5 * j, (4, 4))
    m[0, 0], m[1, 0], m[2, 0], m[3, 0], m[2, 2] = 14, -2, 7, -10, 30
    b = np.array([2.0, -10.0, 5.0, -5.0])
    return m, b


def division_example() -> tuple[np.ndarray, np.ndarray]:
    """examples/Division.cpp:8-16: A(i,j) = 6i - 2j^2 + 2, diag += 5;
    b(i) = 3 + 2i."""
    A = np.fromfunction(lambda i, j: 6.0 * i - 2 * j * j + 2.0, (4, 4))
    A[np.diag_indices(4)] += 5.0
    b = 3.0 + 2.0 * np.arange(4)
    return A, b


def inverse_example() -> tuple[np.ndarray, np.ndarray]:
    """examples/Division.cpp:39-62: the division example's A with
    A(0,0) = 20 (:39), copied to A2 (:48) whose row 0 is then doubled (:61);
    KNOWN["inverse_example"] holds A2.inverse() (:76-83), 5./A2
    (:112-119), A2.inverse() * b (:124-126) and b * A2.inverse() (:131-133)."""
    A, b = division_example()
    A[0, 0] = 20.0
    A[0] *= 2.0
    return A, b


def matrix_example() -> tuple[np.ndarray, np.ndarray]:
    """examples/Matrix.cpp: m1 (4x3) as printed just before `m1 * m5`
    (:257-262) and the 3x3 m5 (:87-91, row-major list)."""
    m1 = np.array(KNOWN["matrix_example"]["m1"], dtype=np.float64)
    m5 = np.array(KNOWN["matrix_example"]["m5"], dtype=np.float64)
    return m1, m5


def decomp_matrix(mattype: int, N: int = 200) -> np.ndarray:
    """tests/TMV_TestMatrixDecomp.cpp:29-106 for square types:
    0 = regular, 3 = singular, 4 = singular with bad defects,
    5 = singular and nearly underflowing, 6 = singular and nearly
    overflowing."""
    M = N
    singular = mattype >= 3
    eps = np.finfo(np.float64).eps
    tmin = np.finfo(np.float64).tiny
    tmax = np.finfo(np.float64).max
    m = np.fromfunction(lambda i, j: 2.0 + 4 * i - 5 * j, (M, N))
    if not singular:
        m /= 10 * N
    m[3, 3] = 14
    m[3, 4] = -2
    m[0, 2] = 7
    m[M - 1, N - 4] = 23
    m[M - 1, N - 2] = 13
    m[M - 1, N - 1] = -10
    if singular:
        m[:, 1] = 0
        m[7, :] = m[6, :]
    else:
        m[np.diag_indices(N)] *= 30 * N
    if mattype == 4:
        for i in range(10, N, 10):
            m[:, i:N] *= 1e-10
            m[i + 5:N, :] *= -1e-10
    if mattype == 5:
        x = tmin
        m[:, :] = x
        for i in range(1, N):
            m[:, i] /= i
        m[N // 2, N // 2] = x / eps
        for i in range(N // 5, 2 * N // 5):        # diag(0, N/5, 2N/5)
            m[i, i] = 0
        for i in range(N // 2 + 1, N):             # diag(0, N/2+1, N) *= -1
            m[i, i] *= -1
        m[N - 4:N, N - 4:N] = -x / eps / 8
    if mattype == 6:
        x = tmax / N
        m[:, :] = x
        for i in range(1, N):
            m[:, i] /= i
        for i in range(N // 5, 2 * N // 5):
            m[i, i] = 0
        for i in range(N // 2, N):
            m[i, i] *= -1
    return m


def big_div_100() -> np.ndarray:
    """tests/TMV_TestMatrixDiv.cpp:155-191 (real part): M = P ^ Q, diag +=
    215, then the pinned row/element/sub-vector edits."""
    k = KNOWN["big_div"]
    P = np.array(k["P"], dtype=np.float64)
    Q = np.array(k["Q"], dtype=np.float64)
    N = len(P)
    M = np.outer(P, Q)
    M[np.diag_indices(N)] += 215
    M[int(np.floor(0.23 * N)), :] *= 12
    M[int(np.floor(0.12 * N)), 1] -= 142
    # subVector(i, j, di, dj, size): start (i,j), step (di,dj), `size` items.
    i, j, s = int(np.floor(0.65 * N)), int(np.floor(0.05 * N)), int(np.floor(0.29 * N))
    for q in range(s):
        M[i + q, j + 3 * q] *= 2
    i, j, s = int(np.floor(0.98 * N)), int(np.floor(0.12 * N)), int(np.floor(0.18 * N))
    for q in range(s):
        M[i - q, j + 2 * q] += 197
    return M
