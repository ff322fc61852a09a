"""The interchange list -> permutation conversion inside tmvx_getrs_*,
through the C ABI: with LU = I the solve returns B with its rows permuted,
so B = 0, 1, ..., n-1 gives the composed permutation itself, which must
equal the oracle's in-order application of the swaps (TMV permuteRows,
src/TMV_Matrix.cpp:866, as LU_LDivEq applies it, src/TMV_LUDiv.cpp:203;
oracle.apply_perm_rows) exactly, and for trans = 1 its inverse.  The
lists cover what the sort-based conversion has to get right: uniform
targets, one target for every step (a single long run),
p_k = k + 1 (a parent chain of length n), half the steps without a swap,
targets clustered near k; a list with some p_k < k (not an LU interchange
list: steps taken in order by one thread); and n past the LDS limit."""
import numpy as np
import pytest
import torch

import tmv_amd

pytestmark = pytest.mark.gpu
SEED = 0x9E12


def lu_list(kind, n, rs):
    k = np.arange(n)
    if kind == "uniform":
        p = rs.integers(k, n)
    elif kind == "one_target":
        p = np.full(n, n - 1)
    elif kind == "chain":
        p = np.minimum(k + 1, n - 1)
    elif kind == "half_self":
        p = np.where(rs.random(n) < 0.5, k, rs.integers(k, n))
    elif kind == "clustered":
        p = np.minimum(k + rs.integers(0, 4, n), n - 1)
    elif kind == "general":                           # some p_k < k
        p = rs.integers(0, n, n)
    else:
        raise ValueError(kind)
    return p.astype(np.int64)


def perm_through_getrs(piv, trans):
    n = len(piv)
    dtype = torch.float64 if n <= 4096 else torch.float32   # 0..n-1 exact either way
    LU = torch.eye(n, dtype=dtype, device="cuda").t()       # column-major identity
    b = torch.arange(n, dtype=dtype, device="cuda")
    tmv_amd.getrs(LU, torch.from_numpy(piv).cuda(), b, trans=trans)
    return b.cpu().numpy().astype(np.int64)


KINDS = ["uniform", "one_target", "chain", "half_self", "clustered", "general"]


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4097, 16384, 32768])
@pytest.mark.parametrize("kind", KINDS)
def test_getrs_permutation_equals_in_order_swaps(orc, dev, kind, n):
    print(f"perm {kind} n={n}", flush=True)
    rs = np.random.default_rng(SEED + n)
    piv = lu_list(kind, n, rs)
    expect = orc.apply_perm_rows(np.arange(n), piv)
    got = perm_through_getrs(piv, trans=False)
    bad = np.nonzero(got != expect)[0]
    assert bad.size == 0, f"first differing row {bad[0]} of {n}"
    if n <= 16384:
        inv = np.empty(n, dtype=np.int64)
        inv[expect] = np.arange(n)
        assert np.array_equal(perm_through_getrs(piv, trans=True), inv)


@pytest.mark.parametrize("kind", ["uniform", "chain"])
def test_getrs_permutation_past_lds_limit(orc, dev, kind):
    """n = 32769: the conversion takes the steps in order in one thread."""
    n = 32769
    print(f"perm {kind} n={n}", flush=True)
    piv = lu_list(kind, n, np.random.default_rng(SEED))
    expect = orc.apply_perm_rows(np.arange(n), piv)
    assert np.array_equal(perm_through_getrs(piv, trans=False), expect)
