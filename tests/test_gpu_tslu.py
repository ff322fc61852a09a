"""Tournament-pivoted FP32 LU (the mixed-precision solver's factorisation),
through the C ABI (tmvx_getrf_f32_tournament).  Not TMV's pivot sequence by
design, so the bar is the factorisation itself: piv is a valid interchange
list (j <= piv[j] < n), ||P A - L U|| <= 64 n eps32 ||A|| (evaluated in FP64
with BLAS: exact enough for an FP32 factorisation, and fast at n = 4100),
|L| entries bounded (no growth blow-up), and a zero column is reported as a
zero pivot."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS32
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def residual64(A, LU, P):
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(P):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return float(np.abs(PA - L @ U).max() / np.abs(A).max())


@pytest.mark.parametrize("n", [1, 5, 64, 100, 513, 1000, 2048, 4100])
def test_tournament_lu(dev, n):
    print(f"tournament getrf n={n}", flush=True)
    A = rng.uniform_matrix(SEED + 700 + n, n, n).astype(np.float32)
    dA = tmv_amd.colmajor(torch.from_numpy(A).cuda())
    piv, info = tmv_amd.getrf_tournament(dA)
    torch.cuda.synchronize()
    assert info == 0
    P = piv.cpu().numpy()
    assert np.all(P >= np.arange(n)) and np.all(P < n)
    LU = dA.cpu().numpy()
    assert residual64(A.astype(np.float64), LU.astype(np.float64), P) <= 64 * n * EPS32
    assert np.abs(np.tril(LU, -1)).max(initial=0.0) <= 64.0


def test_tournament_lu_zero_column(dev):
    n = 700
    A = rng.uniform_matrix(SEED + 7, n, n).astype(np.float32)
    A[:, 300] = 0.0
    dA = tmv_amd.colmajor(torch.from_numpy(A).cuda())
    piv, info = tmv_amd.getrf_tournament(dA)
    torch.cuda.synchronize()
    assert info == 301
    LU = dA.cpu().numpy()
    assert np.isfinite(LU).all()
