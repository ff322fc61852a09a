"""tmvx_getrf_f64_noalloc (no device allocation; scratch in a static device
array) against tmvx_getrf_f64, through the C ABI: two calls in a row, the
second with a smaller matrix, must give bit-identical pivots and factors
and the same info.  The entry point is asynchronous and writes info to a
device word."""
import numpy as np
import pytest
import torch

import tmv_amd
from devutil import to_dev
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def reference(A):
    Ad = to_dev(A)
    piv, info = tmv_amd.getrf(Ad)
    torch.cuda.synchronize()
    return Ad.cpu().numpy(), piv.cpu().numpy(), info


def noalloc(A):
    n = A.shape[0]
    Ad = to_dev(A)
    piv = torch.full((max(n, 1),), -7, dtype=torch.int64, device="cuda")
    d_info = torch.full((1,), -99, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = tmv_amd.lib().tmvx_getrf_f64_noalloc(n, Ad.data_ptr(), max(n, 1), piv.data_ptr(),
                                              d_info.data_ptr(), stream)
    assert rc == tmv_amd.OK, tmv_amd.status_str(rc)
    torch.cuda.synchronize()
    return Ad.cpu().numpy(), piv.cpu().numpy()[:n], int(d_info.item())


@pytest.mark.parametrize("sizes", [(700, 300), (1500, 64), (1030, 1029)])
def test_noalloc_matches_getrf_twice(dev, sizes):
    for k, n in enumerate(sizes):
        print(f"noalloc call {k} n={n}", flush=True)
        A = rng.uniform_matrix(SEED + 70 + n, n, n)
        LU0, P0, info0 = reference(A)
        LU1, P1, info1 = noalloc(A)
        assert np.array_equal(P1, P0)
        assert np.array_equal(LU1, LU0)          # same kernels, same order
        assert info1 == info0 == 0


def test_noalloc_zero_pivot_info(dev):
    """A column that vanishes (TMV's zero-column rule) shows in d_info as in
    the synchronous entry point's info."""
    n = 200
    A = rng.uniform_matrix(SEED + 5, n, n)
    A[:, 37] = 0.0
    LU0, P0, info0 = reference(A)
    LU1, P1, info1 = noalloc(A)
    assert info0 == 38 and info1 == info0
    assert np.array_equal(P1, P0) and np.array_equal(LU1, LU0)


def test_noalloc_rejects_beyond_scratch(dev):
    """n above the scratch array's limit (65536) is refused on the host,
    before any device work (the small buffers are never touched)."""
    A = torch.zeros(16, dtype=torch.float64, device="cuda")
    piv = torch.zeros(16, dtype=torch.int64, device="cuda")
    d_info = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = tmv_amd.lib().tmvx_getrf_f64_noalloc(65537, A.data_ptr(), 65537, piv.data_ptr(),
                                              d_info.data_ptr(), None)
    assert rc == tmv_amd.ERR_UNSUPPORTED
