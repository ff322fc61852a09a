"""GPU parity on the points where index work must be bit-exact and on the
remaining public seams, through the C ABI:

- the whole n = 8192 pivot sequence of the plain FP64 configuration at scale
  equals TMV's (tests/golden: tie-free, min gap > 1e3 * 64 n eps);
- exact ties (row-shuffled Hadamard matrices): pivots and factors equal
  TMV's bit for bit, in FP64 and FP32 (first index wins,
  src/TMV_Vector.cpp:473-504);
- tmvx_trsm_* for every triangle/transpose/unit combination against the
  oracle's TriLDivEq restatement, and its singular report;
- tmvx_laswp_* in both directions against TMV's permuteRows /
  reversePermuteRows restated (apply_perm_rows);
- getrs with trans = 1 and 17 or 64 right-hand sides (the blocked
  multi-RHS path) against the oracle's LU_RDivEq and the backward error;
- the few-RHS solve (nrhs <= 16: reciprocal of the diagonal, one rounding
  more than TMV's division) and the many-RHS solve (nrhs > 16: divides, as
  TMV) on the same system, both within one stated bound of the oracle."""
import hashlib

import numpy as np
import pytest
import torch

import make_ref_vectors as M
import tmv_amd
from checks import EPS64, EPS32, backward_error
from devutil import to_dev, to_host
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def test_cfg2_pivots_equal_tmv(dev, ref_vectors):
    n = 8192
    print(f"getrf n={n} pivots vs TMV", flush=True)
    A = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(A, SEED + 50)
    A = A.view(n, n).t()                       # column-major view of the flat fill
    piv, info = tmv_amd.getrf(A)
    assert info == 0
    P = piv.cpu().numpy()
    Pt = ref_vectors["cfg2_lu_P"]
    bad = np.nonzero(P != Pt)[0]
    assert bad.size == 0, f"first differing step {bad[0]} of {n}"


@pytest.mark.parametrize("n", M.HADAMARD_SIZES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_hadamard_exact_ties(dev, ref_vectors, n, dtype):
    print(f"hadamard ties n={n} {dtype}", flush=True)
    H = M.hadamard_rows(n)
    Ad = to_dev(H, dtype)
    piv, info = tmv_amd.getrf(Ad)
    assert info == 0
    assert np.array_equal(piv.cpu().numpy(), ref_vectors[f"hadamard_{n}_P"])
    LU = to_host(Ad)                            # exact values: FP32 holds them too
    assert sha(LU) == str(ref_vectors[f"hadamard_{n}_LU_sha256"])


def tri(n, upper, seed, zero_diag=None):
    A = rng.uniform_matrix(seed, n, n) / max(n, 1)
    d = 1.0 + rng.uniform(seed + 1, n, 0.0, 1.0)
    A[np.diag_indices(n)] = d * np.where(rng.uniform(seed + 2, n) < 0.5, -1.0, 1.0)
    A = np.triu(A) if upper else np.tril(A)
    if zero_diag is not None:
        A[zero_diag, zero_diag] = 0.0
    return A


@pytest.mark.parametrize("upper", [0, 1])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("unit", [0, 1])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (130, 1), (300, 17), (513, 40)])
def test_trsm_variants(orc, dev, upper, trans, unit, n, nrhs):
    print(f"trsm upper={upper} trans={trans} unit={unit} n={n} nrhs={nrhs}", flush=True)
    A = tri(n, upper, SEED + 730 + n + 4 * upper + 2 * trans + unit)
    B = rng.uniform_matrix(SEED + 740 + n, n, nrhs)
    Bd = to_dev(B)
    tmv_amd.trsm(to_dev(A), Bd, upper=upper, trans=trans, unit=unit)
    X = to_host(Bd)
    Xo = orc.trsm(A, B, upper=upper, trans=trans, unit=unit)
    # well-conditioned triangles (|diag| in [1, 2], off-diagonal <= 1/n):
    # both solves within 8 n eps of the exact one, relative to max |X|
    assert np.abs(X - Xo).max() <= 16 * n * EPS64 * max(np.abs(Xo).max(), 1.0)
    opA = A.T if trans else A
    Aeff = opA - np.diag(np.diag(opA)) + np.eye(n) if unit else opA
    for j in range(nrhs):
        assert backward_error(Aeff, X[:, j], B[:, j]) <= 8 * n * EPS64


@pytest.mark.parametrize("trans", [0, 1])
def test_trsm_singular_reported(orc, dev, trans):
    """A zero diagonal met by a non-zero entry: TMVX_ERR_SINGULAR (TMV throws
    SingularUpperTriMatrix); met only by zeros: a normal solve."""
    n = 200
    A = tri(n, 1, SEED + 750, zero_diag=77)
    B = rng.uniform_matrix(SEED + 751, n, 3)
    with pytest.raises(tmv_amd.SingularError):
        tmv_amd.trsm(to_dev(A), to_dev(B), upper=1, trans=trans)
    b = np.zeros(n)
    if trans:
        b[100:] = 1.0          # A^T lower: rows >= 100 never reach row 77
    else:
        b[:50] = 1.0           # upper: rows < 50 never reach row 77
    bd = to_dev(b)
    tmv_amd.trsm(to_dev(A), bd, upper=1, trans=trans)
    x = to_host(bd)
    xo = orc.trsm(A, b, upper=1, trans=trans)
    assert np.abs(x - xo).max() <= 16 * n * EPS64 * max(np.abs(xo).max(), 1.0)


@pytest.mark.parametrize("kind", ["lu", "general"])
@pytest.mark.parametrize("m,ncols", [(1, 1), (70, 5), (1000, 33)])
def test_laswp_both_directions(orc, dev, kind, m, ncols):
    """Forward: TMV permuteRows (interchanges in order); reverse:
    reversePermuteRows, the inverse (src/TMV_Matrix.cpp:865-930)."""
    g = np.random.default_rng(m + ncols)
    if kind == "lu":                                    # an LU interchange list
        P = np.array([g.integers(i, m) for i in range(m)], np.int64)
    else:                                               # any rows, repeats allowed
        P = g.integers(0, m, m).astype(np.int64)
    A = rng.uniform_matrix(SEED + 760 + m, m, ncols)
    Ad = to_dev(A)
    Pd = torch.from_numpy(P).cuda()
    tmv_amd.laswp(Ad, Pd, forward=True)
    F = to_host(Ad)
    assert np.array_equal(F, orc.apply_perm_rows(A, P))
    tmv_amd.laswp(Ad, Pd, forward=False)
    assert np.array_equal(to_host(Ad), A)               # reverse undoes forward
    # a sub-range k1..k2 of the list
    k1, k2 = m // 3, (2 * m) // 3
    Ad = to_dev(A)
    tmv_amd.laswp(Ad, Pd, k1=k1, k2=k2, forward=True)
    ref = A.copy()
    for i in range(k1, k2):
        if P[i] != i:
            ref[[i, P[i]]] = ref[[P[i], i]]
    assert np.array_equal(to_host(Ad), ref)


@pytest.mark.parametrize("nrhs", [17, 64])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_getrs_trans_many_rhs(orc, dev, nrhs, dtype):
    """A^T X = B with more than 16 right-hand sides (blocked path, trans)."""
    n = 700
    eps = EPS64 if dtype == torch.float64 else EPS32
    A = rng.uniform_matrix(SEED + 770, n, n)
    B = rng.uniform_matrix(SEED + 771, n, nrhs)
    Ad = to_dev(A, dtype)
    piv, info = tmv_amd.getrf(Ad)
    Bd = to_dev(B, dtype)
    tmv_amd.getrs(Ad, piv, Bd, trans=True)
    X = to_host(Bd)
    At = to_host(to_dev(A, dtype)).T                    # A and B as stored
    Bs = to_host(to_dev(B, dtype))
    for j in range(nrhs):
        assert backward_error(At, X[:, j], Bs[:, j]) <= 8 * n * eps
    if dtype == torch.float64:
        LUo, Po = orc.getrf(A)
        Xo = orc.getrs(LUo, Po, B, trans=True)
        assert np.abs(X - Xo).max() <= 1e-9 * np.abs(Xo).max()


def test_few_and_many_rhs_paths_one_bound(orc, dev):
    """The same system solved with 16 right-hand sides (single-launch
    solve: multiplies by the reciprocal of the diagonal) and with 17 (the
    blocked solve: divides, as TMV's LU_LDivEq) -- both within one stated
    bound of the oracle (TMV's order, division): |x - x_oracle| <=
    32 n eps max|x_oracle| per column, and their shared columns agree to
    the same bound."""
    n = 600
    A = rng.uniform_matrix(SEED + 780, n, n) + n * np.eye(n)
    B = rng.uniform_matrix(SEED + 781, n, 17)
    Ad = to_dev(A)
    piv, _ = tmv_amd.getrf(Ad)
    LUo, Po = orc.getrf(A)
    Xo = orc.getrs(LUo, Po, B)
    X16d, X17d = to_dev(B[:, :16]), to_dev(B)
    tmv_amd.getrs(Ad, piv, X16d)
    tmv_amd.getrs(Ad, piv, X17d)
    X16, X17 = to_host(X16d), to_host(X17d)
    bound = 32 * n * EPS64
    for j in range(16):
        s = max(np.abs(Xo[:, j]).max(), 1e-300)
        assert np.abs(X16[:, j] - Xo[:, j]).max() <= bound * s
        assert np.abs(X17[:, j] - Xo[:, j]).max() <= bound * s
        assert np.abs(X16[:, j] - X17[:, j]).max() <= bound * s
    assert np.abs(X17[:, 16] - Xo[:, 16]).max() <= bound * np.abs(Xo[:, 16]).max()
