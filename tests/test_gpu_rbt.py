"""LU without interchanges and the mixed solver's random butterfly
transformation (RBT) on the GPU, through the C ABI.

- tmvx_getrf_nopiv_*: on strictly column-diagonally-dominant matrices
  partial pivoting makes no interchanges, so the oracle's GEPP factors
  (TMV's LU_Decompose restated) are the reference: identity pivots, factors
  within rounding, ||A - LU|| <= 8 n eps ||A||.  An exactly zero pivot is
  reported 1-based.
- tmvx_gesv_mixed_f64 with TMVX_PIVOT_RBT: backward error <= 1e-12 and
  the factorisation it used; a matrix built as U^-T A'' V^-1 from the dense
  butterfly definition (tests/rbt_ref.py) with A''(0,0) = 0 must make the
  device's transformed matrix break down and fall back to partial pivoting,
  which pins the device butterflies to the definition.
"""
import numpy as np
import pytest
import torch

import rbt_ref as R
import tmv_amd
from checks import EPS64, EPS32, backward_error
from devutil import to_dev, to_host
from test_gpu_mixed import conditioned
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x5B7A


def dominant(n, seed):
    return rng.uniform_matrix(seed, n, n) + 2 * n * np.eye(n)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 300, 700])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_getrf_nopiv_matches_gepp_on_dominant(orc, dev, n, dtype):
    print(f"getrf_nopiv n={n} {dtype}", flush=True)
    A = dominant(n, SEED + n)
    LUo, Po = orc.getrf(A)
    assert np.array_equal(Po, np.arange(n))          # GEPP: no interchanges
    Ad = to_dev(A, dtype)
    piv, info = tmv_amd.getrf_nopiv(Ad)
    assert info == 0
    assert np.array_equal(piv.cpu().numpy(), np.arange(n))
    LU = to_host(Ad)
    eps = EPS64 if dtype == torch.float64 else EPS32
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    res = np.abs(A - L @ U).max() / np.abs(A).max()
    assert res <= 8 * n * eps, res
    scale = np.abs(LUo).max()
    assert np.abs(LU - LUo).max() <= 8 * n * eps * scale


def test_getrf_nopiv_reports_zero_pivot(dev):
    """Exact zero pivots: at step 0, and at step 70 (second panel) of a matrix
    whose elimination before it is exact."""
    A = np.eye(5)
    A[0, 0] = 0.0
    A[0, 1] = A[1, 0] = 1.0
    _, info = tmv_amd.getrf_nopiv(to_dev(A))
    assert info == 1
    n = 150
    A = np.eye(n)
    A[70, 70] = 0.0
    A[70, 71] = A[71, 70] = 1.0
    for dtype in (torch.float64, torch.float32):
        _, info = tmv_amd.getrf_nopiv(to_dev(A, dtype))
        assert info == 71


RBT = tmv_amd.LOWPREC_F32 | tmv_amd.PIVOT_RBT


@pytest.mark.parametrize("n,kappa", [(256, 1e2), (257, 1e2), (258, 1e2), (259, 1e2),
                                     (1000, 1e2), (256, 1e6), (259, 1e6), (1000, 1e6)])
def test_mixed_rbt_converges(orc, dev, n, kappa):
    """RBT + no interchanges, n of every residue mod 4 (the identity
    padding), against the FP64 oracle.  Where the growth of the
    interchange-free LU is too large for FP32 (kappa = 1e6 at n = 1000) the
    solve falls back to partial pivoting."""
    print(f"mixed rbt n={n} kappa={kappa:g}", flush=True)
    A = conditioned(n, kappa, SEED + 10 * n)
    B = rng.normal_matrix(SEED + 11 * n, n, 2)
    X, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(B), tol=1e-12, max_iter=40, lowprec=RBT)
    X = to_host(X)
    expect = "rbt->partial" if (n, kappa) == (1000, 1e6) else "rbt"
    assert info["converged"] == 1 and info["pivoting"] == expect, info
    Xo = orc.gesv(A, B)
    for j in range(2):
        assert backward_error(A, X[:, j], B[:, j]) <= 1e-12
        fe = np.abs(X[:, j] - Xo[:, j]).max() / np.abs(Xo[:, j]).max()
        assert fe <= 4 * kappa * max(1e-12, 8 * n * EPS64), (j, fe)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13])
def test_mixed_rbt_tiny(orc, dev, n):
    """Orders below and around one butterfly group (n padded to 4 or 8)."""
    A = rng.uniform_matrix(SEED + 50 + n, n, n) + n * np.eye(n)
    b = rng.uniform(SEED + 60 + n, n, -1.0, 1.0)
    x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b), tol=1e-12, max_iter=40, lowprec=RBT)
    assert info["converged"] == 1 and info["pivoting"].startswith("rbt"), info
    np.testing.assert_allclose(to_host(x), orc.gesv(A, b), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [256, 512])
def test_mixed_rbt_breakdown_falls_back(dev, n):
    """A = U^-T A'' V^-1 with A''(0,0) = 0: the transformed matrix's first
    pivot is zero (or rounding-small), the RBT attempt is abandoned and
    partial pivoting solves the system."""
    App = np.eye(n) + 0.1 * rng.uniform_matrix(SEED + 7, n, n)
    App[0, 0] = 0.0
    U = R.recursive_butterfly(n, 0)
    V = R.recursive_butterfly(n, 1)
    A = np.linalg.solve(U.T, np.linalg.solve(V.T, App.T).T)        # U^-T A'' V^-1
    np.testing.assert_allclose(U.T @ A @ V, App, atol=1e-12)
    b = rng.normal(SEED + 8, n)
    x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b), tol=1e-12, max_iter=40, lowprec=RBT)
    assert info["pivoting"] == "rbt->partial", info
    assert info["converged"] == 1
    assert backward_error(A, to_host(x), b) <= 1e-12


def test_mixed_partial_flag(orc, dev):
    """Partial pivoting is the default and the explicit flag."""
    n = 700
    A = conditioned(n, 1e4, SEED + 21)
    b = rng.normal(SEED + 22, n)
    for lp in (tmv_amd.LOWPREC_F32, tmv_amd.LOWPREC_F32 | tmv_amd.PIVOT_PARTIAL):
        x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b), lowprec=lp)
        assert info["converged"] == 1 and info["pivoting"] == "partial", info
        assert backward_error(A, to_host(x), b) <= 1e-12


def test_mixed_pivot_flags_rejected(dev):
    A = to_dev(np.eye(8))
    b = to_dev(np.ones(8))
    x, info = tmv_amd.gesv_mixed(A, b, check=False,
                                 lowprec=tmv_amd.PIVOT_PARTIAL | tmv_amd.PIVOT_RBT)
    assert info["status"] == tmv_amd.ERR_UNSUPPORTED


@pytest.mark.parametrize("prec", [tmv_amd.LOWPREC_F16, tmv_amd.LOWPREC_BF16])
def test_mixed_rbt_16bit_gmres(dev, prec):
    """16-bit trailing updates after the RBT, GMRES-IR."""
    n = 1024
    A = conditioned(n, 1e2, SEED + 31)
    b = rng.normal(SEED + 32, n)
    x, info = tmv_amd.gesv_mixed(to_dev(A), to_dev(b), lowprec=prec | tmv_amd.PIVOT_RBT,
                                 max_iter=10)
    assert info["converged"] == 1 and info["pivoting"].startswith("rbt"), info
    assert backward_error(A, to_host(x), b) <= 1e-12


def test_mixed_rbt_n16384(dev):
    """The mixed-precision configuration's size, uniform [-1, 1] entries
    generated on the device: the RBT attempt falls back (the growth of the
    interchange-free LU at this size), the result still meets the target;
    backward error recomputed with torch in FP64."""
    n = 16384
    print("mixed rbt n=16384", flush=True)
    A = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(A, SEED + 41)
    A = A.view(n, n).t()                                  # column-major view
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_normal(b, SEED + 42)
    x, info = tmv_amd.gesv_mixed(A, b, tol=1e-12, max_iter=30, lowprec=RBT)
    assert info["converged"] == 1 and info["pivoting"].startswith("rbt"), info
    r = b - A @ x
    be = (r.abs().max() / (A.abs().sum(dim=1).max() * x.abs().max() + b.abs().max())).item()
    assert be <= 1e-12, be
