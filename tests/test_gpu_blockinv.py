"""The mixed solver's correction solve (tmvx_trsm_blockinv_*: 64 x 64
diagonal blocks applied through their inverses, formed off the solve's
dependency chain), through the C ABI, against the oracle's TriLDivEq
restatement (src/TMV_TriDiv_M.cpp:190-266):

- every triangle / transpose / unit-diagonal combination, n with partial
  last blocks and right-hand-side counts spanning two launches, on
  well-conditioned triangles: the substitution test's bounds (16 n eps of
  max |X| forward, 8 n eps backward);
- the factors of a GEPP LU of a random matrix (n = 700): backward error
  within 8 n eps times the largest condition number of the 64 x 64
  diagonal blocks (the explicit inverses' error grows with it), forward
  error within the substitution's times that same factor;
- a zero diagonal is reported as singular."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS64, EPS32, backward_error
from devutil import to_dev, to_host
from test_gpu_parity import tri
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0xB10C


def op_unit(A, trans, unit):
    opA = A.T if trans else A
    return opA - np.diag(np.diag(opA)) + np.eye(len(A)) if unit else opA


@pytest.mark.parametrize("upper", [0, 1])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("unit", [0, 1])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 1), (64, 2), (65, 1), (130, 3), (1000, 5)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_blockinv_variants(orc, dev, upper, trans, unit, n, nrhs, dtype):
    print(f"blockinv upper={upper} trans={trans} unit={unit} n={n} nrhs={nrhs} {dtype}",
          flush=True)
    A = tri(n, upper, SEED + n + 4 * upper + 2 * trans + unit)
    B = rng.uniform_matrix(SEED + 10 + n, n, nrhs)
    Bd = to_dev(B, dtype)
    tmv_amd.trsm_blockinv(to_dev(A, dtype), Bd, upper=upper, trans=trans, unit=unit)
    X = to_host(Bd)
    Xo = orc.trsm(A, B, upper=upper, trans=trans, unit=unit)
    eps = EPS64 if dtype == torch.float64 else EPS32
    assert np.abs(X - Xo).max() <= 16 * n * eps * max(np.abs(Xo).max(), 1.0)
    Aeff = op_unit(A, trans, unit)
    for j in range(nrhs):
        assert backward_error(Aeff, X[:, j], B[:, j]) <= 8 * n * eps


def block_cond(T, nb=64):
    return max(np.linalg.cond(T[b:b + nb, b:b + nb]) for b in range(0, len(T), nb))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_blockinv_on_lu_factors(orc, dev, dtype):
    n = 700
    A = rng.uniform_matrix(SEED + 77, n, n)
    LU, _ = orc.getrf(A)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    b = rng.normal(SEED + 78, n)
    eps = EPS64 if dtype == torch.float64 else EPS32
    LUd = to_dev(LU, dtype)
    for upper, T in ((0, L), (1, U)):
        for trans in (0, 1):
            kb = block_cond(T.T if trans else T)
            bd = to_dev(b, dtype)
            tmv_amd.trsm_blockinv(LUd, bd, upper=upper, trans=trans, unit=not upper)
            x = to_host(bd)
            opT = T.T if trans else T
            be = backward_error(opT, x, b)
            print(f"lu factor upper={upper} trans={trans} block cond {kb:.3g} be {be:.3g}",
                  flush=True)
            assert be <= 8 * n * eps * kb
            xo = orc.trsm(LU, b[:, None], upper=upper, trans=trans, unit=not upper)[:, 0]
            fe = np.abs(x - xo).max() / np.abs(xo).max()
            assert fe <= 16 * n * eps * kb * np.linalg.cond(opT), fe


@pytest.mark.parametrize("upper", [0, 1])
def test_blockinv_zero_diagonal_is_singular(dev, upper):
    n = 130
    A = tri(n, upper, SEED + 5, zero_diag=70)
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.trsm_blockinv(to_dev(A), to_dev(np.ones(n)), upper=upper)
