"""Argument checks of the Python mirror (tmv_amd/__init__.py), on the CPU:
every shape, dtype, pivot or device mismatch is rejected with ValueError /
TypeError before anything reaches the library, so no mismatched buffer is
ever handed to a kernel."""
import pytest
import torch

import tmv_amd

f64, f32 = torch.float64, torch.float32


def cm(m, n, dtype=f64):
    return torch.zeros(n, m, dtype=dtype).t()


def piv(n, dtype=torch.int64):
    return torch.zeros(n, dtype=dtype)


@pytest.mark.parametrize("case", [
    "rows", "dtype", "piv_dtype", "piv_short", "piv_2d", "lu_rect", "b_rowstride"])
def test_getrs_rejects(case):
    LU, P, B = cm(8, 8), piv(8), cm(8, 3)
    exc = ValueError
    if case == "rows":
        B = cm(7, 3)
    elif case == "dtype":
        B, exc = cm(8, 3, f32), TypeError
    elif case == "piv_dtype":
        P, exc = piv(8, torch.int32), TypeError
    elif case == "piv_short":
        P = piv(7)
    elif case == "piv_2d":
        P = torch.zeros(8, 1, dtype=torch.int64)
    elif case == "lu_rect":
        LU = cm(8, 7)
    elif case == "b_rowstride":
        B = torch.zeros(8, 2, dtype=f64)[:, :1]       # stride (2, 1): not a column
    with pytest.raises(exc):
        tmv_amd.getrs(LU, P, B)


def test_host_tensors_rejected():
    """Consistent arguments on the host are still refused: the library reads
    device pointers only."""
    with pytest.raises(ValueError, match="device tensors only"):
        tmv_amd.getrs(cm(8, 8), piv(8), cm(8, 2))
    with pytest.raises(ValueError, match="device tensors only"):
        tmv_amd.gemm(cm(4, 5), cm(5, 6))
    with pytest.raises(ValueError, match="device tensors only"):
        tmv_amd.getrf(cm(4, 4))


@pytest.mark.parametrize("case", ["c_shape", "c_dtype", "inner", "b_dtype", "not2d"])
def test_gemm_rejects(case):
    A, B, C = cm(4, 5), cm(5, 6), cm(4, 6)
    exc = ValueError
    if case == "c_shape":
        C = cm(4, 5)
    elif case == "c_dtype":
        C, exc = cm(4, 6, f32), TypeError
    elif case == "inner":
        B = cm(4, 6)
    elif case == "b_dtype":
        B, exc = cm(5, 6, f32), TypeError
    elif case == "not2d":
        A = torch.zeros(20, dtype=f64)
    with pytest.raises(exc):
        tmv_amd.gemm(A, B, C)


def test_getrf_rejects_rectangular_and_int():
    with pytest.raises(ValueError):
        tmv_amd.getrf(cm(4, 5))
    with pytest.raises(TypeError):
        tmv_amd.getrf(torch.zeros(4, 4, dtype=torch.int32).t())


@pytest.mark.parametrize("fn", ["getri", "logdet"])
def test_factor_users_reject_bad_piv(fn):
    with pytest.raises(TypeError):
        getattr(tmv_amd, fn)(cm(6, 6), piv(6, torch.int32))
    with pytest.raises(ValueError):
        getattr(tmv_amd, fn)(cm(6, 6), piv(5))


def test_trsm_and_laswp_reject():
    with pytest.raises(ValueError):
        tmv_amd.trsm(cm(6, 6), cm(5, 2), upper=True)
    with pytest.raises(TypeError):
        tmv_amd.trsm(cm(6, 6), cm(6, 2, f32), upper=True)
    with pytest.raises(ValueError):
        tmv_amd.laswp(cm(6, 3), piv(6), k1=2, k2=7)
    with pytest.raises(TypeError):
        tmv_amd.laswp(cm(6, 3), piv(6, torch.int32))


def test_mixed_and_adjoint_reject():
    with pytest.raises(TypeError):
        tmv_amd.gesv_mixed(cm(6, 6, f32), torch.zeros(6, dtype=f64))
    with pytest.raises(ValueError):
        tmv_amd.gesv_mixed(cm(6, 6), torch.zeros(5, dtype=f64))
    with pytest.raises(ValueError):
        tmv_amd.gesv_mixed(cm(6, 6), cm(6, 2), X=cm(6, 3))
    with pytest.raises(ValueError):
        tmv_amd.gesv_adjoint(cm(6, 6), piv(6), cm(6, 2), cm(6, 3))
    with pytest.raises(ValueError):
        tmv_amd.gesv_adjoint(cm(6, 6), piv(6), cm(6, 2), cm(6, 2), gradA=cm(5, 5))


def test_fill_rejects():
    with pytest.raises(TypeError):
        tmv_amd.fill_uniform(torch.zeros(4, dtype=f32), 1)
    with pytest.raises(ValueError):
        tmv_amd.fill_normal(torch.zeros(4, dtype=f64), 1)
