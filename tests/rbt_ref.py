"""Random butterfly transformation, restated densely from its definition
for the tests (not the device's group-wise form).

A butterfly of order m is B = 2^-1/2 [R0 R1; R0 -R1] with R0, R1 diagonal of
order m/2; the depth-two recursive butterfly of order N (n rounded up to a
multiple of 4) is W = diag(B1, B2) B, B of order N, B1 and B2 of order N/2
(Parker 1995; Baboulin, Dongarra, Herrmann, Tomov, ACM TOMS 39(2), 2013).
The mixed solver transforms A into U^T A V, U with diagonals 0-5 and V with
diagonals 6-11 of the stream below (tmv_amd/csrc/tmvx_mixed.hip,
rbt_entry / rbt_table_kernel): entry i of diagonal d is exp(u / 10), u =
SplitMix64(seed + golden * ((d << 40) + i + 1)) / 2^64 - 1/2 (53 bits).
"""
import numpy as np

RBT_SEED = 0x5EED5EED0BADCAFE
_M64 = (1 << 64) - 1


def _splitmix(z):
    z &= _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def entries(seed, diag, count):
    out = np.empty(count)
    for i in range(count):
        z = _splitmix(seed + 0x9E3779B97F4A7C15 * ((diag << 40) + i + 1))
        u = (z >> 11) * 2.0 ** -53 - 0.5
        out[i] = np.exp(0.1 * u)
    return out


def butterfly(r0, r1):
    """2^-1/2 [R0 R1; R0 -R1]."""
    h = len(r0)
    B = np.zeros((2 * h, 2 * h))
    i = np.arange(h)
    B[i, i] = r0
    B[i, h + i] = r1
    B[h + i, i] = r0
    B[h + i, h + i] = -r1
    return B / np.sqrt(2.0)


def recursive_butterfly(n, which, seed=RBT_SEED):
    """Dense W (N x N) for U (which = 0) or V (which = 1)."""
    N = 4 * ((n + 3) // 4)
    H, Q = N // 2, N // 4
    d0 = 6 * which
    B = butterfly(entries(seed, d0 + 0, H), entries(seed, d0 + 1, H))
    B1 = butterfly(entries(seed, d0 + 2, Q), entries(seed, d0 + 3, Q))
    B2 = butterfly(entries(seed, d0 + 4, Q), entries(seed, d0 + 5, Q))
    D = np.zeros((N, N))
    D[:H, :H] = B1
    D[H:, H:] = B2
    return D @ B


def group_matrix(n, which, k, seed=RBT_SEED):
    """The 4 x 4 matrix the device uses for positions k + a N/4 (a = 0..3),
    written out as in rbt_table_kernel."""
    N = 4 * ((n + 3) // 4)
    Q = N // 4
    d0 = 6 * which
    e = lambda d, i: entries(seed, d0 + d, i + 1)[i]
    r0a, r0b = e(0, k), e(0, k + Q)
    r1a, r1b = e(1, k), e(1, k + Q)
    s0, s1, t0, t1 = e(2, k), e(3, k), e(4, k), e(5, k)
    return 0.5 * np.array([[s0 * r0a, s1 * r0b, s0 * r1a, s1 * r1b],
                           [s0 * r0a, -s1 * r0b, s0 * r1a, -s1 * r1b],
                           [t0 * r0a, t1 * r0b, -t0 * r1a, -t1 * r1b],
                           [t0 * r0a, -t1 * r0b, -t0 * r1a, t1 * r1b]])


def lu_nopiv(A):
    """Doolittle LU without interchanges (right-looking, reciprocal scaling);
    returns the packed factors."""
    LU = np.array(A, dtype=np.float64)
    n = LU.shape[0]
    for k in range(n):
        LU[k + 1:, k] *= 1.0 / LU[k, k]
        LU[k + 1:, k + 1:] -= np.outer(LU[k + 1:, k], LU[k, k + 1:])
    return LU
