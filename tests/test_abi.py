"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import tmv_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(tmvx_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tmv_amd.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    assert set(tmv_amd.declared_symbols()) == declared()


def test_status_strings_host_only():
    assert tmv_amd.status_str(tmv_amd.OK) == "ok"
    assert "singular" in tmv_amd.status_str(tmv_amd.ERR_SINGULAR)
    assert tmv_amd.lib().tmvx_version().decode().startswith("tmvx")


def test_argument_errors_before_any_device_work():
    """Invalid sizes are rejected on the host (LAPACK info < 0 analogue)."""
    lib = tmv_amd.lib()
    assert lib.tmvx_gemm_f64(0, 0, -1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1, None) \
        == tmv_amd.ERR_ARG
    assert lib.tmvx_gemm_f64(2, 0, 1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1, None) \
        == tmv_amd.ERR_ARG
    info = ctypes.c_int64(0)
    assert lib.tmvx_getrf_f64(4, None, 2, None, ctypes.byref(info), None) == tmv_amd.ERR_ARG
