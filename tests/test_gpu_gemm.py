"""GEMM on the GPU through the C ABI vs the oracle (TMV MultMM restated).
Bar: elementwise |C - C_oracle| <= 2 k eps (|A||B| + |beta||C0|)."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS64, EPS32, gemm_bound_ok
from devutil import to_dev, to_host
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED

SHAPES = [(1, 1, 1), (3, 5, 2), (16, 16, 4), (37, 21, 50), (64, 64, 64), (77, 77, 77),
          (128, 128, 128), (130, 65, 97), (257, 129, 300), (512, 512, 512)]


def _operands(m, n, k, ar, br, dtype):
    A = rng.uniform_matrix(SEED + 21, m, k)
    B = rng.uniform_matrix(SEED + 22, k, n)
    Ad = to_dev(A, dtype, "row" if ar else "col")
    Bd = to_dev(B, dtype, "row" if br else "col")
    return A, B, Ad, Bd


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("ar,br", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_parity(orc, dev, m, n, k, ar, br):
    print(f"dgemm m={m} n={n} k={k} layout={ar}{br}", flush=True)
    A, B, Ad, Bd = _operands(m, n, k, ar, br, torch.float64)
    C = to_host(tmv_amd.gemm(Ad, Bd))
    # the stated bound k eps |A||B| against the exact (long-double) product
    Cl = A.astype(np.longdouble) @ B.astype(np.longdouble)
    ok, worst = gemm_bound_ok(C, Cl, A, B, k, EPS64, factor=1.0)
    assert ok, worst
    # and the oracle (TMV's summation order), each within that bound
    Co = orc.gemm(A, B, a_rowmajor=ar, b_rowmajor=br)
    ok, worst = gemm_bound_ok(C, Co, A, B, k, EPS64, factor=2.0)
    assert ok, worst


@pytest.mark.parametrize("m,n,k", [(64, 48, 80), (200, 100, 33), (512, 512, 512)])
def test_dgemm_add_alpha(orc, dev, m, n, k):
    """TMV C += alpha*A*B (MultMM<true>) and beta generally."""
    A, B, Ad, Bd = _operands(m, n, k, 0, 0, torch.float64)
    C0 = rng.uniform_matrix(SEED + 23, m, n)
    for alpha, beta in ((-0.5, 1.0), (2.0, -3.0), (1.0, 0.0)):
        Cd = to_dev(C0)
        tmv_amd.gemm(Ad, Bd, Cd, alpha=alpha, beta=beta)
        C = to_host(Cd)
        Cl = beta * C0.astype(np.longdouble) + alpha * (A.astype(np.longdouble) @ B.astype(np.longdouble))
        ok, worst = gemm_bound_ok(C, Cl, A, B, k, EPS64, 1.0, C0, beta, alpha)
        assert ok, (alpha, beta, worst)


def test_dgemm_beta_zero_ignores_nan_c(dev):
    """beta == 0 does not read C (TMV MultMM<false> overwrites)."""
    A, B, Ad, Bd = _operands(40, 30, 20, 0, 0, torch.float64)
    Cd = to_dev(np.full((40, 30), np.nan))
    tmv_amd.gemm(Ad, Bd, Cd, beta=0.0)
    assert np.isfinite(to_host(Cd)).all()


def test_dgemm_integer_bit_exact(dev):
    A = np.floor(rng.uniform_matrix(SEED + 24, 300, 200, -50, 50))
    B = np.floor(rng.uniform_matrix(SEED + 25, 200, 260, -50, 50))
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    for ar in (0, 1):
        for br in (0, 1):
            C = to_host(tmv_amd.gemm(to_dev(A, layout="row" if ar else "col"),
                                     to_dev(B, layout="row" if br else "col")))
            assert np.array_equal(C, exact)


def test_dgemm_matches_tmv_cfg1(dev, ref_vectors):
    """512^3: within the elementwise bound of TMV's own product, checked
    through TMV's column sums (stored from the real library)."""
    n = 512
    A = rng.uniform_matrix(SEED + 0, n, n)
    B = rng.uniform_matrix(SEED + 1, n, n)
    C = to_host(tmv_amd.gemm(to_dev(A), to_dev(B)))
    cs = ref_vectors["cfg1_gemm_colsum"]
    bound = 2 * n * EPS64 * (np.abs(A) @ np.abs(B)).sum(axis=0) + n * n * EPS64
    assert np.all(np.abs(C.sum(axis=0) - cs) <= bound)


@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (130, 65, 97), (512, 256, 384)])
@pytest.mark.parametrize("ar,br", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_sgemm_parity(orc, dev, m, n, k, ar, br):
    A, B, Ad, Bd = _operands(m, n, k, ar, br, torch.float32)
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    C = to_host(tmv_amd.gemm(Ad, Bd))
    A64, B64 = A32.astype(np.float64), B32.astype(np.float64)
    Cl = A64.astype(np.longdouble) @ B64.astype(np.longdouble)
    ok, worst = gemm_bound_ok(C, Cl, A64, B64, k, EPS32, factor=1.0)
    assert ok, worst
    Co = orc.gemm(A32, B32, a_rowmajor=ar, b_rowmajor=br, dtype=np.float32)
    ok, worst = gemm_bound_ok(C, Co, A64, B64, k, EPS32, factor=2.0)
    assert ok, worst


@pytest.mark.parametrize("layout", ["CC", "CR", "RC"])
def test_dgemm_8192_checksums(dev, layout):
    """Full size (M=N=K=8192): size-independent checksum properties
    e^T C = (e^T A) B and C e = A (B e), evaluated on the host in float64
    with a rounding bound."""
    n = 8192
    torch.manual_seed(0)
    A = torch.empty(n * n, dtype=torch.float64, device="cuda")
    B = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(A, SEED + 30)
    tmv_amd.fill_uniform(B, SEED + 31)
    # the flat buffer viewed column-major (stride (1, n)) or row-major
    Ad = A.view(n, n).t() if layout[0] == "C" else A.view(n, n)
    Bd = B.view(n, n).t() if layout[1] == "C" else B.view(n, n)
    C = tmv_amd.gemm(Ad, Bd)
    torch.cuda.synchronize()
    Ah, Bh, Ch = to_host(Ad), to_host(Bd), to_host(C)
    e = np.ones(n)
    lhs = Ch @ e
    rhs = Ah @ (Bh @ e)
    bound = 4 * n * EPS64 * (np.abs(Ah) @ (np.abs(Bh) @ e))
    assert np.all(np.abs(lhs - rhs) <= bound)
    lhs = e @ Ch
    rhs = (e @ Ah) @ Bh
    bound = 4 * n * EPS64 * ((e @ np.abs(Ah)) @ np.abs(Bh))
    assert np.all(np.abs(lhs - rhs) <= bound)


@pytest.mark.parametrize("ar,br", [(0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gemm_large_rowmajor_operands(dev, ar, br, dtype):
    """Products past 1024^3 with a row-major operand copy that operand to
    column-major first: the result equals the column-major product of the
    same values bit for bit (same MFMA arithmetic), with alpha and beta, on
    ragged sizes; and it meets the k eps |A||B| bound."""
    m, n, k = 1031, 1043, 1057
    print(f"gemm large {dtype} layout={ar}{br}", flush=True)
    A, B, Ad, Bd = _operands(m, n, k, ar, br, dtype)
    C0 = rng.uniform_matrix(SEED + 23, m, n)
    C = to_dev(C0, dtype)
    tmv_amd.gemm(Ad, Bd, C, alpha=-0.75, beta=0.5)
    Cc = to_dev(C0, dtype)
    tmv_amd.gemm(to_dev(A, dtype), to_dev(B, dtype), Cc, alpha=-0.75, beta=0.5)
    torch.cuda.synchronize()
    assert torch.equal(C, Cc)
    eps = EPS64 if dtype == torch.float64 else EPS32
    Af = A.astype(np.float32) if dtype == torch.float32 else A
    Bf = B.astype(np.float32) if dtype == torch.float32 else B
    C0f = C0.astype(np.float32) if dtype == torch.float32 else C0
    Cl = -0.75 * (Af.astype(np.longdouble) @ Bf.astype(np.longdouble)) + 0.5 * C0f
    mag = 0.75 * (np.abs(Af).astype(np.float64) @ np.abs(Bf).astype(np.float64)) + \
        0.5 * np.abs(C0f)
    err = np.abs(to_host(C) - np.asarray(Cl, np.float64))
    assert (err <= k * eps * mag + 1e-30).all()
