"""LU factorisation and solves on the GPU through the C ABI vs the oracle
(TMV LU_Decompose / LU_LDivEq restated).  Bar: identical pivots on
tie-free inputs, ||P A - L U|| <= 8 n eps ||A||, solve backward error
<= 8 n eps."""
import ctypes

import numpy as np
import pytest
import torch

import fixtures as F
import tmv_amd
from checks import EPS64, EPS32, lu_residual, lu_residual_mn, backward_error
from devutil import to_dev, to_host
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def gpu_getrf(A, dtype=torch.float64):
    Ad = to_dev(A, dtype)
    piv, info = tmv_amd.getrf(Ad)
    return Ad, to_host(Ad), piv.cpu().numpy(), info


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 63, 64, 65, 100, 130, 257, 512, 600, 1000])
def test_getrf_pivots_and_residual(orc, dev, n):
    print(f"getrf n={n}", flush=True)
    A = rng.uniform_matrix(SEED + 400 + n, n, n)
    LUo, Po, gaps = orc.getrf(A, gaps=True)
    _, LU, P, info = gpu_getrf(A)
    assert info == 0
    ok = gaps > 64 * n * EPS64
    k0 = n if ok.all() else int(np.argmin(ok))
    assert np.array_equal(P[:k0], Po[:k0])
    if n <= 600:
        assert lu_residual(A, LU, P) <= 8 * n * EPS64


@pytest.mark.parametrize("mattype", [0, 3, 4, 5, 6])
def test_getrf_tmv_decomp_fixtures(orc, dev, mattype):
    """TMV's own decomposition fixtures, incl. the singular, near-underflow
    (zero-column rule) and near-overflow matrices."""
    A = F.decomp_matrix(mattype)
    LUo, Po, gaps = orc.getrf(A, gaps=True)
    _, LU, P, info = gpu_getrf(A)
    amb = np.nonzero(gaps <= 64 * 200 * EPS64)[0]
    k0 = int(amb[0]) if len(amb) else 200
    assert np.array_equal(P[:k0], Po[:k0])
    if mattype in (0, 3, 4):
        assert lu_residual(A, LU, P) <= 8 * 200 * EPS64
    if mattype == 0:
        assert info == 0
    assert np.isfinite(LU).all()


def test_getrf_zero_column_rule(orc, dev):
    """A column whose candidates underflow is zeroed from the diagonal down
    and interchanged with itself (src/TMV_LUDecompose.cpp:162-169)."""
    n = 70
    A = rng.uniform_matrix(SEED + 41, n, n)
    A[:, 5] = 1e-320          # subnormal: below DBL_MIN
    LUo, Po = orc.getrf(A)
    _, LU, P, info = gpu_getrf(A)
    assert np.array_equal(P, Po)
    assert P[5] == 5 and info == 6
    assert np.all(LU[5:, 5] == 0) and np.all(LUo[5:, 5] == 0)


def test_det_fixtures_gpu(orc, dev):
    mats = F.det_matrices()
    for name, (true, k) in F.KNOWN["dets"].items():
        m = mats[name]
        _, LU, P, _ = gpu_getrf(m)
        d = orc.det_from_lu(LU, P)
        norm_m = mats["m11"] if name == "m12" else m
        assert abs(d - true) <= k * float(np.sum(norm_m ** 2)) * 10 * EPS64, name


@pytest.mark.parametrize("n,nrhs", [(1, 1), (4, 1), (33, 3), (100, 1), (257, 16), (512, 1),
                                    (1000, 16), (300, 2), (200, 6), (130, 7), (3000, 1)])
@pytest.mark.parametrize("trans", [False, True])
def test_getrs(orc, dev, n, nrhs, trans):
    print(f"getrs n={n} nrhs={nrhs} trans={trans}", flush=True)
    A = rng.uniform_matrix(SEED + 500 + n, n, n)
    B = rng.uniform_matrix(SEED + 600 + n, n, nrhs)
    Ad, LU, P, info = gpu_getrf(A)
    pivd = torch.from_numpy(P).cuda()
    Bd = to_dev(B)
    tmv_amd.getrs(Ad, pivd, Bd, trans=trans)
    X = to_host(Bd)
    At = A.T if trans else A
    for j in range(nrhs):
        assert backward_error(At, X[:, j], B[:, j]) <= 8 * n * EPS64
    Xo = orc.getrs(LU, P, B, trans=trans)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=1e4 * n * EPS64 * np.abs(Xo).max())


def test_division_example_gpu(dev):
    A, b = F.division_example()
    Ad, LU, P, _ = gpu_getrf(A)
    bd = to_dev(b)
    tmv_amd.getrs(Ad, torch.from_numpy(P).cuda(), bd)
    np.testing.assert_allclose(to_host(bd), F.KNOWN["division_example"]["x"], rtol=5e-6)


def test_cfg1_solve_matches_tmv(dev, ref_vectors):
    """x = b/A for A + 512 I (the CPU-runnable configuration): pivots equal
    to TMV's and x within 8 n eps backward error and close to TMV's x."""
    n = 512
    A = rng.uniform_matrix(SEED + 0, n, n) + n * np.eye(n)
    b = rng.uniform(SEED + 2, n, -1.0, 1.0)
    Ad, LU, P, _ = gpu_getrf(A)
    assert np.array_equal(P, ref_vectors["cfg1_lu_P"])
    bd = to_dev(b)
    tmv_amd.getrs(Ad, torch.from_numpy(P).cuda(), bd)
    x = to_host(bd)
    assert backward_error(A, x, b) <= 8 * n * EPS64
    np.testing.assert_allclose(x, ref_vectors["cfg1_x"], rtol=0, atol=1e-14)


def test_singular_solve_reports(dev):
    A = F.decomp_matrix(3, N=20)
    Ad, LU, P, info = gpu_getrf(A)
    assert info > 0
    bd = to_dev(np.ones(20))
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.getrs(Ad, torch.from_numpy(P).cuda(), bd)


@pytest.mark.parametrize("n", [64, 300, 1000])
def test_sgetrf(orc, dev, n):
    A = rng.uniform_matrix(SEED + 700 + n, n, n).astype(np.float32)
    LUo, Po, gaps = orc.getrf(A, dtype=np.float32, gaps=True)
    _, LU, P, info = gpu_getrf(A, torch.float32)
    ok = gaps > 64 * n * EPS32
    k0 = n if ok.all() else int(np.argmin(ok))
    assert np.array_equal(P[:k0], Po[:k0])
    if n <= 300:
        assert lu_residual(A, LU.astype(np.float32), P) <= 8 * n * EPS32


@pytest.mark.parametrize("nrhs", [1, 2, 3, 5])
@pytest.mark.parametrize("trans", [False, True])
def test_sgetrs(dev, nrhs, trans):
    """FP32 solve (the refinement's inner solve): backward error in FP32."""
    n = 1000
    A = rng.uniform_matrix(SEED + 710, n, n).astype(np.float32)
    B = rng.uniform_matrix(SEED + 711, n, nrhs).astype(np.float32)
    Ad, LU, P, info = gpu_getrf(A, torch.float32)
    Bd = to_dev(B, torch.float32)
    tmv_amd.getrs(Ad, torch.from_numpy(P).cuda(), Bd, trans=trans)
    X = to_host(Bd).astype(np.float64)
    At = (A.T if trans else A).astype(np.float64)
    for j in range(nrhs):
        assert backward_error(At, X[:, j], B[:, j].astype(np.float64)) <= 8 * n * EPS32


def test_getrf_8192_backward_error(dev):
    """Full size n = 8192 (uniform [-1,1], no shift, 16 RHS): solve
    backward error <= 8 n eps, evaluated on the host in float64."""
    n, nrhs = 8192, 16
    A = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(A, SEED + 50)
    A = A.view(n, n).t()
    B = torch.empty(nrhs * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_normal(B, SEED + 51)
    B = B.view(nrhs, n).t()
    A0 = to_host(A)
    B0 = to_host(B)
    piv, info = tmv_amd.getrf(A)
    assert info == 0
    tmv_amd.getrs(A, piv, B)
    X = to_host(B)
    R = B0 - A0 @ X
    nA = np.abs(A0).sum(axis=1).max()
    be = np.abs(R).max(axis=0) / (nA * np.abs(X).max(axis=0) + np.abs(B0).max(axis=0))
    assert np.all(be <= 8 * n * EPS64), be.max()


def _swapped(piv, b):
    """Rows of b after the interchanges piv[0], piv[1], ... in order."""
    b = b.copy()
    for k, p in enumerate(piv):
        if 0 <= p < len(b) and p != k:
            b[[k, p]] = b[[p, k]]
    return b


@pytest.mark.parametrize("n,kind", [(1, "far"), (63, "far"), (5000, "far"), (5000, "near"),
                                    (5000, "shared"), (700, "backward")])
def test_getrs_interchange_lists(dev, n, kind):
    """getrs with unit factors (L = U = I) applies exactly the interchange
    list: rows of B swapped k <-> piv[k] in order, for lists with far,
    near and shared targets, and (not an LU list) some piv[k] < k."""
    print(f"getrs interchanges n={n} {kind}", flush=True)
    r = np.random.default_rng(SEED + n)
    if kind == "far":
        piv = np.array([r.integers(k, n) for k in range(n)])
    elif kind == "near":
        piv = np.minimum(np.arange(n) + r.integers(0, 3, n), n - 1)
    elif kind == "shared":
        tg = r.integers(0, n, 7)
        piv = np.maximum(np.arange(n), tg[r.integers(0, 7, n)])
    else:
        piv = r.integers(0, n, n)
    B = np.arange(n, dtype=np.float64).reshape(n, 1) + 0.5
    X = to_dev(B)
    tmv_amd.getrs(to_dev(np.eye(n)), torch.from_numpy(piv.astype(np.int64)).cuda(), X)
    assert np.array_equal(to_host(X), _swapped(piv, B))


RECT = [(5, 3), (3, 5), (100, 37), (37, 100), (300, 130), (130, 300), (1, 7), (7, 1)]


@pytest.mark.parametrize("m,n", RECT)
def test_getrf_mn_matches_tmv(orc, dev, ref_vectors, m, n):
    """LU_Decompose of a non-square matrix (tmvx_getrf_mn_f64): the pivots
    are TMV's (fixture written by TMV itself) up to the first near tie,
    P A = L U to 8 max(m, n) eps, and the first min(m, n) columns of the
    factors are TMV's to the same bound.  Past column m of a wide matrix
    the factors are L^-1 P A, as TMV's dgetrf seam gives them; TMV's native
    path leaves those columns off (its P A - L U is O(1) there)."""
    print(f"getrf_mn {m}x{n}", flush=True)
    A = rng.uniform_matrix(SEED + 700 + m + 3 * n, m, n)
    _, _, gaps = orc.getrf(A, gaps=True)
    Ad = to_dev(A)
    piv, info = tmv_amd.getrf_mn(Ad)
    P, LU = piv.cpu().numpy(), to_host(Ad)
    r = min(m, n)
    assert info == 0 and P.shape == (r,)
    ok = gaps > 64 * max(m, n) * EPS64
    k0 = r if ok.all() else int(np.argmin(ok))
    assert np.array_equal(P[:k0], ref_vectors[f"lurect_{m}x{n}_P"][:k0])
    bound = 8 * max(m, n) * EPS64
    assert lu_residual_mn(A, LU, P) <= bound
    if k0 == r:
        LUt = ref_vectors[f"lurect_{m}x{n}_LU"][:, :r]
        assert np.abs(LU[:, :r] - LUt).max() <= bound * max(1.0, np.abs(LUt).max())


@pytest.mark.parametrize("m,n", [(2000, 700), (700, 2000), (1030, 1029), (64, 4000)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_getrf_mn_large(orc, dev, m, n, dtype):
    """Larger tall and wide matrices (several outer panels, the side stream,
    columns past m taking interchanges and U only) against the oracle."""
    print(f"getrf_mn {m}x{n} {dtype}", flush=True)
    A = rng.uniform_matrix(SEED + 710 + m + n, m, n)
    eps = EPS64 if dtype == torch.float64 else EPS32
    Ao = A if dtype == torch.float64 else A.astype(np.float32)
    _, Po, gaps = orc.getrf(Ao, dtype=np.float64 if dtype == torch.float64 else np.float32,
                             gaps=True)
    Ad = to_dev(Ao, dtype)
    piv, info = tmv_amd.getrf_mn(Ad)
    P, LU = piv.cpu().numpy(), to_host(Ad)
    assert info == 0
    ok = gaps > 64 * max(m, n) * eps
    k0 = min(m, n) if ok.all() else int(np.argmin(ok))
    if dtype == torch.float64:            # FP32 gaps fall below 64 n eps32 early
        assert k0 > min(m, n) // 4
    assert np.array_equal(P[:k0], Po[:k0])
    assert lu_residual_mn(Ao, LU, P) <= 8 * max(m, n) * eps


def test_getrf_mn_square_is_getrf(dev):
    """m == n through the rectangular entry is the square factorisation,
    bit for bit (the same kernels in the same order)."""
    n = 700
    A = rng.uniform_matrix(SEED + 720, n, n)
    A1, A2 = to_dev(A), to_dev(A)
    p1, _ = tmv_amd.getrf(A1)
    p2, _ = tmv_amd.getrf_mn(A2)
    assert torch.equal(p1, p2) and torch.equal(A1, A2)


def test_getrf_mn_empty_and_args(dev):
    for (m, n) in ((0, 5), (5, 0), (0, 0)):
        piv, info = tmv_amd.getrf_mn(tmv_amd.colmajor(torch.zeros(m, n, dtype=torch.float64).cuda()))
        assert info == 0 and piv.numel() == 0
    info = ctypes.c_int64(0)                                  # lda < m
    assert tmv_amd.lib().tmvx_getrf_mn_f64(5, 3, None, 4, None, ctypes.byref(info),
                                           None) == tmv_amd.ERR_ARG


def test_getrf_n40000_f32(dev):
    """n = 40000 (past the 32768 limit of earlier versions): FP32 LU and a
    solve, backward error <= 8 n eps32 in FP64 arithmetic on the device."""
    n = 40000
    print(f"getrf f32 n={n}", flush=True)
    A64 = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(A64, SEED + 730)
    A = A64.view(n, n).t().to(torch.float32)                 # column-major FP32
    del A64
    LU = A.clone()
    piv, info = tmv_amd.getrf(LU)
    assert info == 0
    b = torch.ones(n, dtype=torch.float32, device="cuda")
    x = b.clone()
    tmv_amd.getrs(LU, piv, x)
    del LU
    Ad = A.double()
    r = b.double() - Ad @ x.double()
    be = float(r.abs().max() / (Ad.abs().sum(dim=1).max() * x.double().abs().max() + 1.0))
    assert be <= 8 * n * EPS32, be
