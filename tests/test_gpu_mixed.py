"""Mixed-precision solve (FP32 LU on MFMA + FP64 iterative refinement) vs
the FP64 oracle.  Matrices A = U diag(s) V^T with U, V orthogonal factors of
seeded Gaussian matrices and s log-spaced from 1 to 1/kappa (the
mixed-precision configuration, at test sizes).  Bar: normwise backward
error <= 1e-12 (evaluated independently on the host in long double) and
forward error vs the FP64 oracle within 4 * kappa * max(tol, 8 n eps64)
(stopping at backward error tol bounds the forward error by about kappa *
tol)."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS64, backward_error
from devutil import to_dev, to_host
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def conditioned(n, kappa, seed):
    G1 = rng.normal_matrix(seed, n, n)
    G2 = rng.normal_matrix(seed + 1, n, n)
    U, _ = np.linalg.qr(G1)
    V, _ = np.linalg.qr(G2)
    s = np.logspace(0, -np.log10(kappa), n)
    return (U * s) @ V.T


@pytest.mark.parametrize("n", [256, 700, 1500])
@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("nrhs", [1, 3, 20])
def test_mixed_ir(orc, dev, n, kappa, nrhs):
    print(f"mixed n={n} kappa={kappa:g} nrhs={nrhs}", flush=True)
    A = conditioned(n, kappa, SEED + 800 + n)
    B = rng.normal_matrix(SEED + 900 + n, n, nrhs)
    X, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(B), tol=1e-12, max_iter=40)
    X = to_host(X)
    assert info["converged"] == 1, info
    assert info["backward_error"] <= 1e-12
    Xo = orc.gesv(A, B)
    for j in range(nrhs):
        assert backward_error(A, X[:, j], B[:, j]) <= 1e-12
        fe = np.abs(X[:, j] - Xo[:, j]).max() / np.abs(Xo[:, j]).max()
        assert fe <= 4 * kappa * max(1e-12, 8 * n * EPS64), (j, fe)


def test_mixed_uniform_plain_config(orc, dev):
    """The plain configuration's matrix (uniform [-1,1] + n I) converges in
    very few steps."""
    n = 512
    A = rng.uniform_matrix(SEED + 0, n, n) + n * np.eye(n)
    b = rng.uniform(SEED + 2, n, -1.0, 1.0)
    x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b))
    assert info["converged"] and info["iterations"] <= 4
    np.testing.assert_allclose(to_host(x), orc.gesv(A, b), rtol=0, atol=1e-14)


def test_mixed_reports_not_converged(dev):
    """With max_iter = 0 the target is not reachable: status reported, not raised."""
    n = 300
    A = conditioned(n, 1e6, SEED + 77)
    b = rng.normal(SEED + 78, n)
    x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b), max_iter=0, check=False)
    assert info["status"] == tmv_amd.ERR_NOT_CONVERGED and info["converged"] == 0


def test_mixed_singular_reports(dev):
    A = np.ones((64, 64))
    b = np.ones(64)
    x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b), check=False)
    assert info["status"] == tmv_amd.ERR_SINGULAR and info["factor_info"] > 0


@pytest.mark.parametrize("n,nrhs", [(1000, 1), (1000, 4), (1000, 16), (777, 5)])
def test_getrs_single_launch_solve(orc, dev, n, nrhs):
    """getrs with few right-hand sides runs the single-launch blocked solve."""
    A = rng.uniform_matrix(SEED + 950 + n, n, n)
    B = rng.uniform_matrix(SEED + 951 + n, n, nrhs)
    Ad = to_dev(A)
    piv, info = tmv_amd.getrf(Ad)
    for trans in (False, True):
        Bd = to_dev(B)
        tmv_amd.getrs(Ad, piv, Bd, trans=trans)
        X = to_host(Bd)
        At = A.T if trans else A
        for j in range(nrhs):
            assert backward_error(At, X[:, j], B[:, j]) <= 8 * n * EPS64
