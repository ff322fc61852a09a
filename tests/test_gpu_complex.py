"""Complex GEMM (tmvx_gemm_z / tmvx_gemm_c) through the C ABI vs the oracle's
restatement of TMV's complex product (four real products).

Bar: exact on integer-valued parts (every product and sum exact); elsewhere
|C - C_ref| <= 4 k u (|A|_1 |B|_1) |alpha| + 2 u |beta||C0| elementwise,
|.|_1 = |re| + |im| and u the unit roundoff of the real type, in every
layout pair and conjugation; TMV's own output (bit-exact with the oracle on
aligned shapes) to the same bound; torch's complex product at 4096."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS32, EPS64
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def cplx(seed, m, n):
    return rng.uniform_matrix(seed, m, n) + 1j * rng.uniform_matrix(seed + 1, m, n)


def to_dev(z, dtype, rowmajor=False):
    t = torch.from_numpy(np.ascontiguousarray(z)).to(device="cuda", dtype=dtype)
    return t if rowmajor else tmv_amd.colmajor(t)


def abs1(z):
    return np.abs(z.real) + np.abs(z.imag)


def check(Z, ref, A, B, alpha, beta, C0, eps, k):
    bound = 4 * max(k, 1) * eps * abs(alpha) * (abs1(A) @ abs1(B)) + 1e-300
    if C0 is not None and beta != 0:
        bound = bound + 2 * eps * abs(beta) * np.abs(C0)
    err = np.abs(Z - ref)
    return float((err / bound).max())


@pytest.mark.parametrize("la,lb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("ca,cb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_zgemm_vs_oracle(orc, dev, la, lb, ca, cb):
    m, n, k = 133, 70, 97
    print(f"zgemm layouts {la}{lb} conj {ca}{cb}", flush=True)
    A, B, C0 = cplx(SEED + 1, m, k), cplx(SEED + 3, k, n), cplx(SEED + 5, m, n)
    alpha, beta = 0.75 - 0.25j, -0.5 + 1.0j
    dA, dB = to_dev(A, torch.complex128, la), to_dev(B, torch.complex128, lb)
    dC = to_dev(C0, torch.complex128)
    tmv_amd.gemm_complex(dA, dB, dC, alpha=alpha, beta=beta, conj_a=ca, conj_b=cb)
    Z = dC.cpu().numpy()
    # oracle: C = alpha op(A) op(B), then + beta C0
    ref = orc.gemm_complex(A, B, alpha=alpha, conj_a=ca, conj_b=cb) + beta * C0
    assert check(Z, ref, A, B, alpha, beta, C0, EPS64, k) <= 1.0


@pytest.mark.parametrize("ca,cb", [(0, 0), (1, 1)])
def test_cgemm_vs_oracle(orc, dev, ca, cb):
    m, n, k = 100, 130, 257
    A, B = cplx(SEED + 11, m, k), cplx(SEED + 13, k, n)
    A32, B32 = A.astype(np.complex64), B.astype(np.complex64)
    dA, dB = to_dev(A32, torch.complex64), to_dev(B32, torch.complex64)
    Z = tmv_amd.gemm_complex(dA, dB, alpha=1.0 - 2.0j, conj_a=ca, conj_b=cb).cpu().numpy()
    ref = orc.gemm_complex(A32, B32, alpha=1.0 - 2.0j, conj_a=ca, conj_b=cb,
                           dtype=np.complex64).astype(np.complex128)
    assert check(Z.astype(np.complex128), ref, A32, B32, 1.0 - 2.0j, 0, None, EPS32, k) <= 1.0


@pytest.mark.parametrize("ca,cb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_zgemm_integer_exact(dev, ca, cb):
    g = np.random.default_rng(11 + 2 * ca + cb)
    m, n, k = 65, 47, 129
    A = g.integers(-9, 10, (m, k)) + 1j * g.integers(-9, 10, (m, k))
    B = g.integers(-9, 10, (k, n)) + 1j * g.integers(-9, 10, (k, n))
    C0 = g.integers(-9, 10, (m, n)) + 1j * g.integers(-9, 10, (m, n))
    dC = to_dev(C0, torch.complex128)
    tmv_amd.gemm_complex(to_dev(A, torch.complex128), to_dev(B, torch.complex128, True), dC,
                         alpha=2.0 - 3.0j, beta=1.0 + 1.0j, conj_a=ca, conj_b=cb)
    Aop = A.conj() if ca else A
    Bop = B.conj() if cb else B
    assert np.array_equal(dC.cpu().numpy(), (2.0 - 3.0j) * (Aop @ Bop) + (1.0 + 1.0j) * C0)


def test_zgemm_matches_tmv_output(orc, dev, ref_vectors):
    """TMV's ragged-shape output (its small complex kernels) within the bound."""
    A = rng.uniform_matrix(SEED + 121, 37, 41) + 1j * rng.uniform_matrix(SEED + 122, 37, 41)
    B = rng.uniform_matrix(SEED + 123, 41, 29) + 1j * rng.uniform_matrix(SEED + 124, 41, 29)
    Z = tmv_amd.gemm_complex(to_dev(A, torch.complex128), to_dev(B, torch.complex128),
                             alpha=0.3 + 0.7j, conj_a=True).cpu().numpy()
    assert check(Z, ref_vectors["zgemm_37x29x41_conjA"], A, B, 0.3 + 0.7j, 0, None, EPS64,
                 41) <= 1.0


def test_zgemm_edge_sizes(dev):
    # k = 0: C = beta C; m = 0 / n = 0: nothing to do
    C0 = cplx(SEED + 31, 5, 4)
    dC = to_dev(C0, torch.complex128)
    dA = torch.zeros(5, 0, dtype=torch.complex128, device="cuda")
    dB = torch.zeros(0, 4, dtype=torch.complex128, device="cuda")
    tmv_amd.gemm_complex(dA, dB, dC, alpha=1.0, beta=2.0 - 1.0j)
    np.testing.assert_allclose(dC.cpu().numpy(), (2.0 - 1.0j) * C0, rtol=1e-15)
    e = torch.zeros(0, 3, dtype=torch.complex128, device="cuda")
    assert tmv_amd.gemm_complex(e, to_dev(cplx(1, 3, 2), torch.complex128)).shape == (0, 2)
    A, B = cplx(SEED + 33, 1, 1), cplx(SEED + 35, 1, 1)
    Z = tmv_amd.gemm_complex(to_dev(A, torch.complex128), to_dev(B, torch.complex128))
    np.testing.assert_allclose(Z.cpu().numpy(), A @ B, rtol=1e-15)


def test_zgemm_4096_vs_torch(dev):
    """4096^3 complex128 against torch's own complex product (rocBLAS)."""
    n = 4096
    g = torch.Generator(device="cuda").manual_seed(5)
    A = torch.randn(n, n, dtype=torch.complex128, device="cuda", generator=g)
    B = torch.randn(n, n, dtype=torch.complex128, device="cuda", generator=g)
    Z = tmv_amd.gemm_complex(tmv_amd.colmajor(A), tmv_amd.colmajor(B))
    R = A @ B
    mag = (A.real.abs() + A.imag.abs()) @ (B.real.abs() + B.imag.abs())
    assert ((Z - R).abs() <= 8 * n * EPS64 * mag).all()


# ---------------------------------------------------------------- complex LU ----

def zresidual(A, LU, P):
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = np.array(A, np.complex128)
    for i, p in enumerate(P):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return float(np.abs(PA - L @ U).max() / np.abs(A).max())


def gpu_zgetrf(A, dtype=torch.complex128):
    dA = to_dev(A, dtype)
    piv, info = tmv_amd.getrf_complex(dA)
    return dA, dA.cpu().numpy().astype(np.complex128), piv.cpu().numpy(), info


@pytest.mark.parametrize("n", [1, 2, 5, 37, 64, 65, 100, 257, 700])
def test_zgetrf_pivots_and_residual(orc, dev, ref_vectors, n):
    """Pivots equal to the oracle's (and TMV's where the fixtures hold them)
    where the pivot gap exceeds 64 n eps; P A = L U to 8 n eps."""
    print(f"zgetrf n={n}", flush=True)
    if n in (5, 37, 100, 257):
        A = rng.uniform_matrix(SEED + 230 + n, n, n) + 1j * rng.uniform_matrix(SEED + 240 + n, n, n)
    else:
        A = cplx(SEED + 300 + n, n, n)
    LUo, Po, gaps = orc.getrf_complex(A, gaps=True)
    _, LU, P, info = gpu_zgetrf(A)
    assert info == 0
    ok = gaps > 64 * n * EPS64
    k0 = n if ok.all() else int(np.argmin(ok))
    assert np.array_equal(P[:k0], Po[:k0])
    if f"zlu_{n}_P" in ref_vectors:
        assert np.array_equal(P, ref_vectors[f"zlu_{n}_P"])
    assert zresidual(A, LU, P) <= 8 * n * EPS64


def test_zgetrf_zero_column_rule(orc, dev):
    n = 90
    A = cplx(SEED + 401, n, n)
    A[:, 11] = 0.0
    _, LU, P, info = gpu_zgetrf(A)
    _, Po = orc.getrf_complex(A)
    assert info == 12 and P[11] == 11
    assert np.all(LU[11:, 11] == 0)
    assert np.array_equal(P[:11], Po[:11])


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (33, 3), (257, 2), (600, 5)])
def test_zgetrs(orc, dev, n, nrhs, trans):
    print(f"zgetrs n={n} nrhs={nrhs} trans={trans}", flush=True)
    A, B = cplx(SEED + 500 + n, n, n), cplx(SEED + 520 + n, n, nrhs)
    dA, LU, P, _ = gpu_zgetrf(A)
    dB = to_dev(B, torch.complex128)
    tmv_amd.getrs_complex(dA, torch.from_numpy(P).cuda(), dB, trans=trans)
    X = dB.cpu().numpy()
    M = A.T if trans else A
    r = np.abs(B - M @ X).max() / (np.abs(M).max() * np.abs(X).max() * n)
    assert r <= 8 * EPS64
    Xo = orc.getrs_complex(LU, P, B, trans=trans)     # oracle on the GPU's factors
    assert np.abs(X - Xo).max() <= 1e-10 * max(1.0, np.abs(Xo).max())


def test_zgetrs_singular_reports(dev):
    n = 20
    A = cplx(SEED + 611, n, n)
    A[:, 4] = 0.0
    dA, _, P, info = gpu_zgetrf(A)
    assert info == 5
    dB = to_dev(cplx(SEED + 613, n, 1), torch.complex128)
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.getrs_complex(dA, torch.from_numpy(P).cuda(), dB)


def test_cgetrf_complex64(orc, dev):
    n = 300
    A = cplx(SEED + 700, n, n)
    A32 = A.astype(np.complex64)
    _, LU, P, info = gpu_zgetrf(A32, torch.complex64)
    assert info == 0
    assert zresidual(A32.astype(np.complex128), LU, P) <= 8 * n * EPS32
    dA = to_dev(A32, torch.complex64)
    piv, _ = tmv_amd.getrf_complex(dA)
    B = cplx(SEED + 701, n, 1).astype(np.complex64)
    dB = to_dev(B, torch.complex64)
    tmv_amd.getrs_complex(dA, piv, dB)
    X = dB.cpu().numpy().astype(np.complex128)
    r = np.abs(B - A32.astype(np.complex128) @ X).max() / (np.abs(A32).max() * np.abs(X).max() * n)
    assert r <= 8 * EPS32


@pytest.mark.parametrize("n", [1, 3, 37, 300])
def test_zlogdet_and_inverse(orc, dev, n):
    """Complex log|det| and sign against numpy's slogdet (the determinant
    does not depend on the pivot order), the inverse against A X = I and
    the oracle's solve of A X = I on the GPU's factors."""
    A = cplx(SEED + 800 + n, n, n)
    dA, LU, P, info = gpu_zgetrf(A)
    ld, sg = tmv_amd.logdet_complex(dA, torch.from_numpy(P).cuda())
    nsg, nld = np.linalg.slogdet(A)
    assert abs(ld - nld) <= 1e-10 * max(1.0, abs(nld))
    assert abs(sg - nsg) <= 1e-10
    X = tmv_amd.getri_complex(dA, torch.from_numpy(P).cuda()).cpu().numpy()
    r = np.abs(A @ X - np.eye(n)).max() / (n * np.abs(A).max() * np.abs(X).max())
    assert r <= 8 * EPS64
    Xo = orc.getrs_complex(LU, P, np.eye(n, dtype=np.complex128))
    assert np.abs(X - Xo).max() <= 1e-10 * max(1.0, np.abs(Xo).max())


def test_zlogdet_singular(dev):
    A = cplx(SEED + 901, 20, 20)
    A[:, 3] = 0.0
    dA, _, P, info = gpu_zgetrf(A)
    assert info == 4
    ld, sg = tmv_amd.logdet_complex(dA, torch.from_numpy(P).cuda())
    assert ld == -np.inf and sg == 0
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.getri_complex(dA, torch.from_numpy(P).cuda())
