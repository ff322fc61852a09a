"""FP16 / BF16 trailing updates (tmvx_gemm_f32_lowp) and the mixed-precision
solver with FP16 / BF16 factorisations (GMRES-IR by default, classical
refinement on request), through the C ABI.

Bar for the GEMM: elementwise |C - C_ref| <= 2 (2 u16 + k u32) (|A||B|)_ij
(+ beta term) against a long-double product of the unrounded FP32 inputs,
u16 the unit roundoff of the 16-bit format (FP16 2^-11, BF16 2^-8): each
operand is rounded once, products of two 16-bit values are exact in FP32,
sums accumulate in FP32.  Scaling keeps FP16 accurate for B far outside its
range.  For the solver: backward error <= 1e-12 (recomputed on the host in
long double) with GMRES-IR for FP16 up to kappa = 1e4 (n up to 4096) and
BF16 at kappa = 1e2; classical refinement converges at kappa = 1e2.  Where
the 16-bit factors are too weak a preconditioner for GMRES(50) (BF16 at
kappa = 1e4, FP16 at 1e6 on these log-spaced spectra) the solver reports
not-converged instead of returning an unconverged iterate."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS32, backward_error
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED
# operand error per product: one rounding to 16 bits (u16 each), or, split
# into hi + 2^-S lo, the rounding of lo (u16^2) plus the dropped lo x lo
# term (u16^2): 3 u16^2 per product, written as 2 u_eff below with u_eff =
# 1.5 u16^2
U16 = {tmv_amd.LOWPREC_F16: 2.0 ** -11, tmv_amd.LOWPREC_BF16: 2.0 ** -8,
       tmv_amd.LOWPREC_F16X2: 1.5 * 2.0 ** -22, tmv_amd.LOWPREC_BF16X2: 1.5 * 2.0 ** -16}
NAME = {tmv_amd.LOWPREC_F16: "f16", tmv_amd.LOWPREC_BF16: "bf16",
        tmv_amd.LOWPREC_F16X2: "f16x2", tmv_amd.LOWPREC_BF16X2: "bf16x2"}


def gemm_check(m, n, k, lp, scale_b=True, bscale=1.0, alpha=-1.0, beta=1.0, seed=0):
    A = rng.uniform_matrix(SEED + seed, m, k).astype(np.float32)
    B = (rng.uniform_matrix(SEED + seed + 1, k, n) * bscale).astype(np.float32)
    C0 = rng.uniform_matrix(SEED + seed + 2, m, n).astype(np.float32)
    dA = tmv_amd.colmajor(torch.from_numpy(A).cuda())
    dB = tmv_amd.colmajor(torch.from_numpy(B).cuda())
    dC = tmv_amd.colmajor(torch.from_numpy(C0).cuda())
    tmv_amd.gemm_lowp(dA, dB, dC, alpha=alpha, beta=beta, lowprec=lp, scale_b=scale_b)
    torch.cuda.synchronize()
    C = dC.cpu().numpy().astype(np.float64)
    ref = alpha * (A.astype(np.longdouble) @ B.astype(np.longdouble)) + beta * C0
    mag = abs(alpha) * (np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64))
    bound = 2 * (2 * U16[lp] + k * EPS32) * mag + 2 * EPS32 * abs(beta) * np.abs(C0) + 1e-30
    err = np.abs(C - np.asarray(ref, np.float64))
    return float((err / bound).max()), C


@pytest.mark.parametrize("lp", [tmv_amd.LOWPREC_F16, tmv_amd.LOWPREC_BF16,
                                tmv_amd.LOWPREC_F16X2, tmv_amd.LOWPREC_BF16X2])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (128, 128, 32), (300, 257, 256),
                                   (1000, 700, 64), (513, 129, 300)])
def test_gemm_lowp_bound(dev, lp, m, n, k):
    print(f"gemm_lowp {NAME[lp]} {m}x{n}x{k}", flush=True)
    worst, _ = gemm_check(m, n, k, lp, seed=m + n + k)
    assert worst <= 1.0, worst


@pytest.mark.parametrize("plain,split", [(tmv_amd.LOWPREC_F16, tmv_amd.LOWPREC_F16X2),
                                         (tmv_amd.LOWPREC_BF16, tmv_amd.LOWPREC_BF16X2)])
def test_gemm_lowp_split_gains_bits(dev, plain, split):
    """The split operands' error is far below the single-term one (a split
    that dropped its second term would fail this)."""
    e1, C1 = gemm_check(300, 257, 256, plain, seed=11)
    e2, C2 = gemm_check(300, 257, 256, split, seed=11)
    A = rng.uniform_matrix(SEED + 11, 300, 256)
    B = rng.uniform_matrix(SEED + 12, 256, 257)
    C0 = rng.uniform_matrix(SEED + 13, 300, 257)
    ref = -(A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)) \
        + C0.astype(np.float32)
    d1 = np.abs(C1 - ref).max()
    d2 = np.abs(C2 - ref).max()
    assert d2 < d1 / 64, (d1, d2)


@pytest.mark.parametrize("bscale", [1e6, 1e-6])
def test_gemm_lowp_f16_scaling(dev, bscale):
    """B far above FP16's range (1e6 > 65504) or deep in its subnormals:
    the per-call power-of-two scale keeps the product accurate."""
    worst, C = gemm_check(256, 256, 256, tmv_amd.LOWPREC_F16, scale_b=True, bscale=bscale,
                          seed=7)
    assert np.isfinite(C).all()
    assert worst <= 1.0, worst


def conditioned(n, kappa, seed):
    g = np.random.default_rng(seed)
    U, _ = np.linalg.qr(g.standard_normal((n, n)))
    V, _ = np.linalg.qr(g.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(kappa), n)
    return (U * s) @ V.T


def conditioned_gpu(n, kappa, seed):
    """The same construction on the GPU (torch QR), column-major FP64."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    U, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64).cuda())
    V, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64).cuda())
    s = torch.logspace(0, -np.log10(kappa), n, dtype=torch.float64, device="cuda")
    return tmv_amd.colmajor((U * s[None, :]) @ V.t().contiguous())


def solve_check(A, dA, b, lp, max_iter=20):
    db = torch.from_numpy(b).cuda()
    x, info = tmv_amd.gesv_mixed(dA, db, tol=1e-12, max_iter=max_iter, lowprec=lp)
    torch.cuda.synchronize()
    assert info["converged"] == 1 and info["status"] == tmv_amd.OK, info
    assert info["backward_error"] <= 1e-12
    assert backward_error(A, x.cpu().numpy(), b) <= 1e-12
    return info


F16, BF16 = tmv_amd.LOWPREC_F16, tmv_amd.LOWPREC_BF16


@pytest.mark.parametrize("lp,n,kappa", [(F16, 600, 1e2), (F16, 2048, 1e2), (F16, 600, 1e4),
                                        (F16, 1500, 1e4), (BF16, 600, 1e2), (BF16, 2048, 1e2)])
def test_mixed_lowp_gmres_converges(dev, lp, n, kappa):
    """GMRES-IR (the default after 16-bit trailing updates)."""
    print(f"mixed gmres {NAME[lp]} n={n} kappa={kappa:g}", flush=True)
    A = conditioned(n, kappa, SEED + n + int(np.log10(kappa)))
    b = np.random.default_rng(n).standard_normal(n)
    info = solve_check(A, tmv_amd.colmajor(torch.from_numpy(A).cuda()), b, lp)
    assert info["inner_iterations"] >= info["iterations"] > 0


def test_mixed_lowp_gmres_n4096_kappa1e4(dev):
    """n = 4096, kappa = 1e4: FP16 trailing updates with GMRES-IR reach a
    backward error <= 1e-12."""
    n = 4096
    print(f"mixed gmres f16 n={n} kappa=1e4", flush=True)
    dA = conditioned_gpu(n, 1e4, SEED + 4096)
    A = dA.cpu().numpy()
    b = np.random.default_rng(4096).standard_normal(n)
    solve_check(A, dA, b, F16)


@pytest.mark.parametrize("lp", [tmv_amd.LOWPREC_F16X2, tmv_amd.LOWPREC_BF16X2])
def test_mixed_split_n4096_kappa1e4(dev, lp):
    """n = 4096, kappa = 1e4 with split-operand 16-bit trailing updates
    (classical refinement after FP16X2, GMRES-IR after BF16X2 by default)."""
    n = 4096
    print(f"mixed {NAME[lp]} n={n} kappa=1e4", flush=True)
    dA = conditioned_gpu(n, 1e4, SEED + 4097)
    A = dA.cpu().numpy()
    b = np.random.default_rng(4097).standard_normal(n)
    solve_check(A, dA, b, lp, max_iter=30)


@pytest.mark.parametrize("lp,n,kappa", [(tmv_amd.LOWPREC_F16X2, 600, 1e6),
                                        (tmv_amd.LOWPREC_BF16X2, 600, 1e6),
                                        (tmv_amd.LOWPREC_BF16X2, 1500, 1e4)])
def test_mixed_split_converges(dev, lp, n, kappa):
    A = conditioned(n, kappa, SEED + n + int(np.log10(kappa)))
    b = np.random.default_rng(n).standard_normal(n)
    solve_check(A, tmv_amd.colmajor(torch.from_numpy(A).cuda()), b, lp, max_iter=30)


@pytest.mark.parametrize("lp,kappa", [(BF16, 1e4), (F16, 1e6)])
def test_mixed_lowp_gmres_stall_reported(dev, lp, kappa):
    """Beyond what GMRES(50) can recover from the 16-bit factors (BF16 at
    kappa = 1e4, FP16 at 1e6, n = 600): a converged status only with a
    backward error <= 1e-12, otherwise not-converged, never a wrong answer
    reported as converged."""
    n = 600
    A = conditioned(n, kappa, SEED + n + int(np.log10(kappa)))
    b = np.random.default_rng(n).standard_normal(n)
    dA = tmv_amd.colmajor(torch.from_numpy(A).cuda())
    x, info = tmv_amd.gesv_mixed(dA, torch.from_numpy(b).cuda(), tol=1e-12, max_iter=3,
                                 lowprec=lp, check=False)
    torch.cuda.synchronize()
    be = backward_error(A, x.cpu().numpy(), b)
    if info["status"] == tmv_amd.OK:
        assert info["converged"] == 1 and be <= 1e-12
    else:
        assert info["status"] == tmv_amd.ERR_NOT_CONVERGED and info["converged"] == 0
        assert abs(info["backward_error"] - be) <= 1e-3 * be   # the reported error is the iterate's


def test_mixed_f32_gmres_kappa1e6(dev):
    """GMRES-IR on request after an FP32 factorisation, kappa = 1e6."""
    n = 1000
    A = conditioned(n, 1e6, SEED + 31)
    b = np.random.default_rng(31).standard_normal(n)
    info = solve_check(A, tmv_amd.colmajor(torch.from_numpy(A).cuda()), b,
                       tmv_amd.LOWPREC_F32 | tmv_amd.REFINE_GMRES)
    assert info["inner_iterations"] > 0


@pytest.mark.parametrize("lp", [tmv_amd.LOWPREC_F16, tmv_amd.LOWPREC_BF16])
def test_mixed_lowp_classic_kappa1e2(dev, lp):
    """Classical refinement on request after 16-bit updates (kappa = 1e2)."""
    n = 600
    A = conditioned(n, 1e2, SEED + n)
    b = np.random.default_rng(n).standard_normal(n)
    info = solve_check(A, tmv_amd.colmajor(torch.from_numpy(A).cuda()), b,
                       lp | tmv_amd.REFINE_CLASSIC, max_iter=60)
    assert info["inner_iterations"] == 0


def test_mixed_lowp_reports_not_converged(dev):
    """kappa = 1e6 is beyond what BF16 trailing updates plus classical
    refinement can correct in 5 steps: the solver says so instead of
    returning a wrong answer as converged."""
    n = 1024
    A = conditioned(n, 1e6, SEED + 9)
    b = np.random.default_rng(9).standard_normal(n)
    dA = tmv_amd.colmajor(torch.from_numpy(A).cuda())
    db = torch.from_numpy(b).cuda()
    x, info = tmv_amd.gesv_mixed(dA, db, tol=1e-12, max_iter=5,
                                 lowprec=tmv_amd.LOWPREC_BF16 | tmv_amd.REFINE_CLASSIC,
                                 check=False)
    torch.cuda.synchronize()
    if info["status"] == tmv_amd.OK:
        assert backward_error(A, x.cpu().numpy(), b) <= 1e-12
    else:
        assert info["status"] == tmv_amd.ERR_NOT_CONVERGED and info["converged"] == 0


def test_mixed_refine_flags_rejected(dev):
    """Both refinement flags at once, an unknown flag bit or an unknown
    precision: unsupported."""
    A = tmv_amd.colmajor(torch.eye(8, dtype=torch.float64).cuda())
    b = torch.ones(8, dtype=torch.float64).cuda()
    for lp in (tmv_amd.REFINE_GMRES | tmv_amd.REFINE_CLASSIC, 256, 5):
        x, info = tmv_amd.gesv_mixed(A, b, lowprec=lp, check=False)
        assert info["status"] == tmv_amd.ERR_UNSUPPORTED
