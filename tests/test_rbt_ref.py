"""The random butterfly transformation's restatements against each other and
against the mathematics (CPU only): the device's 4 x 4 group form equals
the dense product of the butterfly definition; W is invertible with
W^-1 = B^-1 diag(B1, B2)^-1 from the butterflies' closed-form inverses; and
x = V (U^T A V)^-1 U^T b solves A x = b through an LU without interchanges."""
import numpy as np
import pytest

import rbt_ref as R
from tmv_amd import rng


@pytest.mark.parametrize("n", [4, 8, 20, 64])
@pytest.mark.parametrize("which", [0, 1])
def test_group_form_is_the_dense_product(n, which):
    W = R.recursive_butterfly(n, which)
    Q = n // 4
    for k in range(Q):
        idx = [k + a * Q for a in range(4)]
        np.testing.assert_allclose(W[np.ix_(idx, idx)], R.group_matrix(n, which, k),
                                   rtol=1e-15, atol=0)
    # W touches nothing outside the groups
    mask = np.zeros((n, n), dtype=bool)
    for k in range(Q):
        idx = [k + a * Q for a in range(4)]
        mask[np.ix_(idx, idx)] = True
    assert np.all(W[~mask] == 0)


def test_butterfly_inverse_closed_form():
    """B = 2^-1/2 [R0 R1; R0 -R1] has B^-1 = 2^-1/2 [R0^-1 R0^-1; R1^-1 -R1^-1]."""
    r0 = R.entries(R.RBT_SEED, 0, 8)
    r1 = R.entries(R.RBT_SEED, 1, 8)
    B = R.butterfly(r0, r1)
    Binv = np.block([[np.diag(1 / r0), np.diag(1 / r0)],
                     [np.diag(1 / r1), -np.diag(1 / r1)]]) / np.sqrt(2.0)
    np.testing.assert_allclose(B @ Binv, np.eye(16), atol=1e-15)


def test_entries_range_and_spread():
    """exp(u/10), u in [-1/2, 1/2): entries in [e^-0.05, e^0.05), not constant."""
    e = R.entries(R.RBT_SEED, 3, 4096)
    assert e.min() >= np.exp(-0.05) and e.max() < np.exp(0.05)
    assert e.std() > 0.02


@pytest.mark.parametrize("n", [32, 100])
def test_rbt_then_lu_without_interchanges_solves(n):
    A = rng.uniform_matrix(0x5B7 + n, n, n)
    b = rng.uniform(0x5B8 + n, n, -1.0, 1.0)
    N = 4 * ((n + 3) // 4)
    Ap = np.eye(N)
    Ap[:n, :n] = A
    U = R.recursive_butterfly(n, 0)
    V = R.recursive_butterfly(n, 1)
    LU = R.lu_nopiv(U.T @ Ap @ V)
    L = np.tril(LU, -1) + np.eye(N)
    Uf = np.triu(LU)
    z = U.T @ np.concatenate([b, np.zeros(N - n)])
    y = np.linalg.solve(Uf, np.linalg.solve(L, z))
    x = (V @ y)[:n]
    np.testing.assert_allclose(A @ x, b, atol=1e-9)


def test_lu_nopiv_matches_gepp_on_dominant_matrix():
    """A strictly column-diagonally-dominant matrix needs no interchanges:
    the restated LU without interchanges equals numpy's GEPP factors."""
    import scipy.linalg
    n = 40
    A = rng.uniform_matrix(0x5B9, n, n) + 2 * n * np.eye(n)
    P, L, U = scipy.linalg.lu(A)
    assert np.array_equal(P, np.eye(n))
    LU = R.lu_nopiv(A)
    np.testing.assert_allclose(np.tril(LU, -1) + np.eye(n), L, atol=1e-14)
    np.testing.assert_allclose(np.triu(LU), U, rtol=1e-13)
