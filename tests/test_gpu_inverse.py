"""Inverse (tmvx_getri_*) and log-determinant (tmvx_logdet_*) from the GPU
factors, through the C ABI, vs the oracle and the real TMV's outputs.

Bar: the inverse matches TMV's m.inverse() to A's conditioning
(||X - X_tmv|| <= 64 n eps kappa ||X_tmv||) and A X = I to a few eps; the
printed values of examples/Division.cpp to 6 significant digits; logdet
to 4 n eps relative and the sign exactly; exact determinants of
tests/TMV_TestMatrixDet.cpp within TMV's own tolerance."""
import numpy as np
import pytest
import torch

import fixtures as F
import tmv_amd
from checks import EPS64, EPS32, inverse_residual, inverse_close
from devutil import to_dev, to_host
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED
EPS = 10 * EPS64          # TMV's test EPS (tests/TMV_Test.h:24)


def factor(A, dtype=torch.float64):
    Ad = to_dev(A, dtype)
    piv, info = tmv_amd.getrf(Ad)
    return Ad, piv, info


def test_inverse_example_golden(dev):
    A2, b = F.inverse_example()
    K = F.KNOWN["inverse_example"]
    LU, piv, info = factor(A2)
    assert info == 0
    X = to_host(tmv_amd.getri(LU, piv))
    np.testing.assert_allclose(X, K["A2inv"], rtol=5e-6)
    np.testing.assert_allclose(5.0 * X, K["five_over_A2"], rtol=5e-6)
    np.testing.assert_allclose(X @ b, K["inv_times_b"], rtol=5e-6)
    np.testing.assert_allclose(b @ X, K["b_times_inv"], rtol=5e-6)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 37, 64, 100, 130])
def test_inverse_matches_tmv(orc, dev, ref_vectors, n):
    print(f"getri n={n}", flush=True)
    M = rng.uniform_matrix(SEED + 200 + n, n, n)
    LU, piv, info = factor(M)
    X = to_host(tmv_amd.getri(LU, piv))
    ok, r = inverse_close(X, ref_vectors[f"inv_{n}"], M, EPS64)
    assert ok, r
    assert inverse_residual(M, X) <= 8 * EPS64


@pytest.mark.parametrize("n", [255, 600, 1000, 2049])
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_inverse_vs_oracle(orc, dev, n, dt):
    """Larger sizes (several blocks of the multi-RHS solves) against the
    oracle's solve of A X = I on the GPU's own factors."""
    print(f"getri n={n} {dt}", flush=True)
    eps = EPS64 if dt == torch.float64 else EPS32
    ndt = np.float64 if dt == torch.float64 else np.float32
    M = rng.uniform_matrix(SEED + 700 + n, n, n)
    Mr = M.astype(ndt).astype(np.float64)
    LU, piv, info = factor(M, dt)
    assert info == 0
    X = to_host(tmv_amd.getri(LU, piv)).astype(np.float64)
    Xo = orc.inverse_from_lu(to_host(LU), piv.cpu().numpy(), dtype=ndt).astype(np.float64)
    ok, r = inverse_close(X, Xo, Mr, eps)
    assert ok, r
    assert inverse_residual(Mr, X) <= 16 * eps


def test_inverse_8192_property(dev):
    """At n = 8192 (beyond what the oracle finishes in seconds): A X - I
    evaluated on the GPU in FP64 stays at a few eps times the pivot growth
    max|U| / max|A| (the factor in the backward error bound of GEPP)."""
    n = 8192
    A = torch.empty(n, n, dtype=torch.float64, device=dev)
    tmv_amd.fill_uniform(A, SEED + 8192)
    A = tmv_amd.colmajor(A)
    A0 = A.clone()
    piv, info = tmv_amd.getrf(A)
    assert info == 0
    X = tmv_amd.getri(A, piv)
    growth = torch.triu(A).abs().max().item() / A0.abs().max().item()
    R = A0 @ X
    R.diagonal().sub_(1.0)
    res = R.abs().max().item() / (n * A0.abs().max().item() * X.abs().max().item())
    print(f"n={n} residual {res / EPS64:.2f} eps, growth {growth:.1f}", flush=True)
    assert res <= 8 * EPS64 * max(1.0, growth), (res, growth)


def test_inverse_singular_reports(dev):
    A = F.decomp_matrix(3, N=20)
    LU, piv, info = factor(A)
    assert info > 0
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.getri(LU, piv)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 37, 64, 100, 130])
def test_logdet_matches_tmv(orc, dev, ref_vectors, n):
    M = rng.uniform_matrix(SEED + 200 + n, n, n)
    LU, piv, _ = factor(M)
    ld, sg = tmv_amd.logdet(LU, piv)
    ref_ld, ref_sg = ref_vectors[f"logdet_{n}"]
    assert sg == ref_sg
    assert abs(ld - ref_ld) <= 4 * n * EPS64 * max(1.0, abs(ref_ld))
    # and equal to the oracle on the same factors (summation order aside)
    old, osg = orc.logdet_from_lu(to_host(LU), piv.cpu().numpy())
    assert osg == sg and abs(ld - old) <= 4 * n * EPS64 * max(1.0, abs(old))


def test_det_fixtures_device(dev):
    mats = F.det_matrices()
    for name, (true, k) in F.KNOWN["dets"].items():
        m = mats[name]
        LU, piv, _ = factor(m)
        d = tmv_amd.det(LU, piv)
        norm_m = mats["m11"] if name == "m12" else m
        assert abs(d - true) <= k * float(np.sum(norm_m ** 2)) * EPS, (name, d, true)


def test_logdet_cfg1_and_singular(dev, ref_vectors):
    n = 512
    As = rng.uniform_matrix(SEED + 0, n, n) + n * np.eye(n)
    LU, piv, _ = factor(As)
    ld, sg = tmv_amd.logdet(LU, piv)
    assert sg == ref_vectors["cfg1_logdet"][1]
    assert abs(ld - ref_vectors["cfg1_logdet"][0]) <= 4 * n * EPS64 * ld
    LU, piv, info = factor(F.decomp_matrix(3, N=20))
    assert info > 0
    assert tmv_amd.logdet(LU, piv) == (-np.inf, 0.0)
    assert tmv_amd.det(LU, piv) == 0.0


@pytest.mark.parametrize("n", [100, 4096])
def test_logdet_f32_and_large(orc, dev, n):
    """FP32 factors (logdet still summed in FP64) and a 4096 FP64 case,
    against the oracle on the same downloaded factors."""
    for dt, ndt in ((torch.float32, np.float32), (torch.float64, np.float64)):
        M = rng.uniform_matrix(SEED + 900 + n, n, n)
        LU, piv, _ = factor(M, dt)
        ld, sg = tmv_amd.logdet(LU, piv)
        old, osg = orc.logdet_from_lu(to_host(LU).astype(np.float64), piv.cpu().numpy())
        assert sg == osg
        assert abs(ld - old) <= 4 * n * EPS64 * max(1.0, abs(old)), (dt, ld, old)


def test_empty_matrix(dev):
    LU = torch.empty(0, 0, dtype=torch.float64, device=dev)
    piv = torch.empty(0, dtype=torch.int64, device=dev)
    assert tmv_amd.logdet(LU, piv) == (0.0, 1.0)
    assert tmv_amd.getri(LU, piv).shape == (0, 0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.complex128,
                                   torch.complex64])
@pytest.mark.parametrize("n", [1, 7, 300, 2049])
def test_lu_diag_range(dev, dtype, n):
    """min / max Abs2(U_ii) (|re| + |im| for complex, TMV's isSingular
    measure, src/TMV_LUD.cpp:195-200) reduced on the device equal the host's
    scan of the same diagonal, exactly (a min / max picks an element)."""
    g = np.random.default_rng(n)
    M = g.uniform(-1, 1, (n, n))
    if dtype.is_complex:
        M = M + 1j * g.uniform(-1, 1, (n, n))
    M[n // 2, n // 2] = 0.0                       # one zero on the diagonal
    d = tmv_amd.colmajor(torch.from_numpy(M).to(dtype).cuda())
    mn, mx = tmv_amd.lu_diag_range(d)
    dg = np.diag(d.cpu().numpy())
    if dtype.is_complex:                          # summed in FP64, as the kernel does
        a = np.abs(dg.real).astype(np.float64) + np.abs(dg.imag).astype(np.float64)
    else:
        a = np.abs(dg).astype(np.float64)
    assert mn == a.min() and mx == a.max()


def test_lu_diag_range_empty(dev):
    mn, mx = tmv_amd.lu_diag_range(tmv_amd.colmajor(torch.empty(0, 0, dtype=torch.float64).cuda()))
    assert mn == float("inf") and mx == 0.0
