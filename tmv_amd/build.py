"""Build libtmvx.so (HIP, gfx950) in-tree.

    python -m tmv_amd.build

Each ``csrc/*.hip`` file is compiled to an object under ``tmv_amd/_build``
in parallel, then linked into ``tmv_amd/libtmvx.so``.  Always recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtmvx.so")
ARCH = os.environ.get("TMVX_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # Packed FP32 instructions stay off (v_pk_* forms are avoided).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def _gen_gemm_instances() -> list[str]:
    """One translation unit per GEMM tile configuration listed in
    csrc/tmvx_gemm_cfgs.h, so the instantiations compile in parallel."""
    import re
    gen = os.path.join(BUILD, "gen")
    os.makedirs(gen, exist_ok=True)
    table = open(os.path.join(CSRC, "tmvx_gemm_cfgs.h")).read()
    out = []
    for dt, macro in (("double", "TMVX_F64_CFGS"), ("float", "TMVX_F32_CFGS")):
        body = table.split("#define " + macro, 1)[1].split("\n\n", 1)[0]
        for m in re.finditer(r"X\((\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", body):
            cid, bm, bn, bk, wm, wn, mw = m.groups()
            cfg = f"GemmCfg<{dt}, {bm}, {bn}, {bk}, {wm}, {wn}, {mw}>"
            src = (f'#include "tmvx_gemm_kernel.h"\nnamespace tmvx {{\n'
                   f"template int gemm_launch_cfg<{cfg}, {dt}>(int, int, i64, i64, i64, {dt}, "
                   f"const {dt}*, i64, const {dt}*, i64, {dt}, {dt}*, i64, hipStream_t);\n}}\n")
            path = os.path.join(gen, f"gemm_{dt}_{cid}.hip")
            if not os.path.exists(path) or open(path).read() != src:
                with open(path, "w") as f:
                    f.write(src)
            out.append(path)
    return out


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_tmv_shim.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_tmv_shim")


def build_cpp_tests() -> str:
    """The C++ program testing the TMV-shaped layer (include/tmvx_tmv.h)
    against libtmvx.so; run by tests/test_gpu_shim.py."""
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC,
           "-o", CPP_TEST_BIN + ".tmp", "-L", HERE, "-ltmvx",
           "-Wl,-rpath,$ORIGIN/../../tmv_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{r.stderr[-6000:]}")
    os.replace(CPP_TEST_BIN + ".tmp", CPP_TEST_BIN)
    return CPP_TEST_BIN


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + _gen_gemm_instances()
    with cf.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, 16, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    build_cpp_tests()
    if verbose:
        print(f"built {LIB} and {CPP_TEST_BIN}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
