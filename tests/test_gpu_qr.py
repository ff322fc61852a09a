"""QR decomposition (tmvx_geqrf_*) and QR division (tmvx_geqrs_*) through the
C ABI vs the oracle's restatement of TMV's QR (and TMV's own outputs).

Bar: beta and R equal to the oracle's within 64 m eps ||A|| (the reflector
convention is TMV's, so no sign freedom is left), A = Q R to 8 m eps ||A||
with Q rebuilt from the GPU's reflectors, least-squares solutions equal to
the oracle's (and TMV's x = b/A fixtures) to 1e-12 on these well-conditioned
inputs, zero columns give beta = 0 and a singular R is reported."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS32, EPS64
from devutil import to_dev, to_host
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def q_from(QR, beta):
    m, n = QR.shape
    Q = np.eye(m)
    for j in reversed(range(n)):
        v = np.concatenate(([1.0], QR[j + 1:, j]))
        Q[j:] -= beta[j] * np.outer(v, v @ Q[j:])
    return Q


def gpu_qr(A, dtype=torch.float64):
    dA = to_dev(A, dtype)
    beta = tmv_amd.geqrf(dA)
    return dA, to_host(dA), beta.cpu().double().numpy(), beta


@pytest.mark.parametrize("m,n", [(1, 1), (5, 3), (37, 37), (100, 64), (300, 257), (1000, 700),
                                 (2000, 1500)])
def test_geqrf_vs_oracle(orc, dev, m, n):
    print(f"geqrf {m}x{n}", flush=True)
    A = rng.uniform_matrix(SEED + 260 + m + n, m, n)
    _, QR, beta, _ = gpu_qr(A)
    QRo, bo = orc.qr_decompose(A)
    tol = 64 * m * EPS64 * np.abs(A).max()
    assert np.abs(beta - bo).max() <= 64 * m * EPS64
    assert np.abs(np.triu(QR[:n]) - np.triu(QRo[:n])).max() <= tol
    assert np.abs(np.tril(QR, -1) - np.tril(QRo, -1)).max() <= 64 * m * EPS64 * 10
    if m <= 1000:
        R = np.zeros((m, n))
        R[:n] = np.triu(QR[:n])
        assert np.abs(A - q_from(QR, beta) @ R).max() <= 8 * m * EPS64 * np.abs(A).max()


@pytest.mark.parametrize("m,n", [(37, 37), (100, 64)])
def test_geqrf_matches_tmv(dev, ref_vectors, m, n):
    A = rng.uniform_matrix(SEED + 260 + m + n, m, n)
    _, QR, beta, _ = gpu_qr(A)
    assert np.abs(beta - ref_vectors[f"qr_{m}x{n}_beta"]).max() <= 64 * m * EPS64
    assert np.abs(np.triu(QR[:n]) - ref_vectors[f"qr_{m}x{n}_R"]).max() <= \
        64 * m * EPS64 * np.abs(A).max()


@pytest.mark.parametrize("m,n,nrhs", [(37, 37, 2), (100, 64, 2), (700, 300, 3), (1500, 1500, 1)])
def test_geqrs_least_squares(orc, dev, ref_vectors, m, n, nrhs):
    print(f"geqrs {m}x{n} nrhs={nrhs}", flush=True)
    A = rng.uniform_matrix(SEED + 260 + m + n, m, n)
    B = rng.uniform_matrix(SEED + 270 + m, m, nrhs)
    dA, QR, beta, dbeta = gpu_qr(A)
    dB = to_dev(B)
    tmv_amd.geqrs(dA, dbeta, dB)
    X = to_host(dB)[:n]
    Xo = orc.qr_solve(QR, beta, B)               # oracle on the GPU's factors
    assert np.abs(X - Xo).max() <= 1e-10 * max(1.0, np.abs(Xo).max())
    if f"qr_{m}x{n}_x" in ref_vectors and nrhs == 2:
        np.testing.assert_allclose(X, ref_vectors[f"qr_{m}x{n}_x"], rtol=0, atol=1e-12)
    if m == n:
        r = np.abs(B - A @ X).max() / (np.abs(A).max() * np.abs(X).max() * n)
        assert r <= 16 * EPS64
    else:
        np.testing.assert_allclose(X, np.linalg.lstsq(A, B, rcond=None)[0], rtol=0, atol=1e-11)


def test_geqrf_zero_column_and_singular(orc, dev):
    A = rng.uniform_matrix(SEED + 281, 40, 12)
    A[:, 5] = 0.0
    dA, QR, beta, dbeta = gpu_qr(A)
    assert QR[5, 5] == 0.0 and beta[5] == 0.0
    QRo, bo = orc.qr_decompose(A)
    assert np.abs(np.triu(QR[:12]) - np.triu(QRo[:12])).max() <= 1e-12
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.geqrs(dA, dbeta, to_dev(np.ones(40)))


def test_geqrf_f32(orc, dev):
    m, n = 400, 300
    A = rng.uniform_matrix(SEED + 290, m, n).astype(np.float32)
    dA, QR, beta, dbeta = gpu_qr(A, torch.float32)
    R = np.zeros((m, n))
    R[:n] = np.triu(QR[:n])
    assert np.abs(A - q_from(QR, beta) @ R).max() <= 8 * m * EPS32 * np.abs(A).max()
    B = rng.uniform_matrix(SEED + 291, m, 1).astype(np.float32)
    dB = to_dev(B, torch.float32)
    tmv_amd.geqrs(dA, dbeta, dB)
    X = to_host(dB)[:n]
    Xl = np.linalg.lstsq(A.astype(np.float64), B.astype(np.float64), rcond=None)[0]
    assert np.abs(X - Xl).max() <= 1e-3 * np.abs(Xl).max()
