"""The TMV-shaped C++ layer (include/tmvx_tmv.h), tested by a C++ program
written like TMV's own tests (tests/cpp/test_tmv_shim.cpp): x = b/m, b%m,
divideUsing / saveDiv / resetDiv semantics, det / logDet, LU_Decompose,
row-major (isTrans) inputs, MultMM into row-major C, tmv::Singular."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_tmv_shim")


@pytest.mark.gpu
def test_tmv_shim_program():
    assert os.path.exists(BIN), "build first: python -m tmv_amd.build"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout


def test_tmv_shim_program_built():
    """build() compiles the C++ test program (no GPU needed to check)."""
    assert os.path.exists(BIN)
