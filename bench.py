#!/usr/bin/env python3
"""Headline benchmark: LU-solve TFLOP/s, FP32-MFMA + iterative refinement vs
plain FP64, at a normwise backward error <= 1e-12; DGEMM % of FP64 peak.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one mixed-precision solve x = b/A of the mixed-precision
configuration: A = U diag(s) V^T, n = 16384, U and V orthogonal factors of
seeded Gaussian matrices (QR), s log-spaced from 1 to 1/kappa (kappa =
1e4), b ~ N(0,1), 1 right-hand side; FP32 LU with partial pivoting on MFMA,
FP64 residual and update on the GPU until backward error <= 1e-12.
`value` = effective FP64-equivalent TFLOP/s, ((2/3) n^3 + 2 n^2) / time per
solve (the HPL-MxP convention), inputs resident in HBM.

Also measured in the same run (outside the K timed steps):
  - the plain FP64 GPU LU solve (getrf + getrs) on the same A and b;
  - DGEMM M = N = K = 8192 in the CCC, CRC and RCC operand layouts; the
    CCC kernel is the `roofline` entry (FP64 MFMA, 78.6 TFLOP/s peak),
    timed with HIP events on the stream it runs on;
  - the CPU baseline: the real TMV 0.76 library (oracle/_ref, built from
    the reference sources) on a bounded sample, x = b/A at n = 4096 with
    OpenMP on the host's cores (kind "reference"); the oracle restatement
    (kind "port") where the TMV build is absent.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
SEED = 0x7A15EED
# HBM bytes per launch of the 8192^3 CCC DGEMM (gemm_kernel f64 128x128x16),
# from rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0), the factors a
# 1 GiB device copy reads on the same counters; profiles/dgemm_8192_ccc_traffic.md.
DGEMM_8192_CCC_TRAFFIC = 13.71e9


def conditioned_gpu(n, kappa, seed, torch, tmv_amd):
    """A = U diag(s) V^T (column-major FP64 on the GPU)."""
    g = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_normal(g, seed)
    U, _ = torch.linalg.qr(g.view(n, n))
    tmv_amd.fill_normal(g, seed + 1)
    V, _ = torch.linalg.qr(g.view(n, n))
    del g
    s = torch.logspace(0, -torch.log10(torch.tensor(kappa)).item(), n,
                       dtype=torch.float64, device="cuda")
    US = (U * s[None, :]).contiguous()          # row-major storage of U diag(s)
    del U
    A = tmv_amd.gemm(US, V.t())                 # column-major U diag(s) V^T
    del US, V
    torch.cuda.empty_cache()
    return A


def backward_error_gpu(A, x, b, torch):
    r = b - A @ x
    nA = A.abs().sum(dim=1).max()
    return float(r.abs().max() / (nA * x.abs().max() + b.abs().max()))


def cpu_baseline(seconds_cap=30.0):
    """TMV's own CPU path on a bounded sample: x = b/A, n = 4096, best of 3."""
    import numpy as np
    import oracle
    from tmv_amd import rng
    n = 4096
    A = rng.uniform_matrix(SEED + 4000, n, n)
    b = rng.uniform(SEED + 4001, n, -1.0, 1.0)
    lib = oracle.ref_lib()
    kind = "reference" if lib is not None else "port"
    if lib is None:
        oracle.build()
    best = None
    t_start = time.perf_counter()
    for _ in range(3):
        t0 = time.perf_counter()
        if lib is not None:
            oracle.ref_gesv(A, b)
        else:
            oracle.gesv(A, b)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        if time.perf_counter() - t_start > seconds_cap:
            break
    flops = 2.0 / 3.0 * n ** 3 + 2.0 * n * n
    threads = (lib or oracle.oracle_lib()).fn("num_threads")()
    return {"value": round(flops / best / 1e12, 5), "unit": "TFLOP/s", "cores": int(threads),
            "kind": kind,
            "sample": (f"x = b/A (divideUsing LU: LU_Decompose + LU_LDivEq, OpenMP MultMM), "
                       f"n={n}, uniform[-1,1] seeded, best of 3, {best * 1e3:.0f} ms"),
            "ms": round(best * 1e3, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--kappa", type=float, default=1e4)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--fp64-reps", type=int, default=2)
    ap.add_argument("--gemm-n", type=int, default=8192)
    ap.add_argument("--gemm-reps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "design is described in DESIGN.md and is not measured", file=sys.stderr)
        sys.exit(2)

    import torch
    import tmv_amd
    tmv_amd.lib()                      # the HIP library; no fallback
    n, nrhs = a.n, a.nrhs

    # ---- inputs (untimed) -------------------------------------------------
    t_setup = time.perf_counter()
    A = conditioned_gpu(n, a.kappa, SEED + 10, torch, tmv_amd)
    b = torch.empty(n * nrhs, dtype=torch.float64, device="cuda")
    tmv_amd.fill_normal(b, SEED + 12)
    b = b.view(nrhs, n).t() if nrhs > 1 else b
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup

    # ---- the timed steps: mixed-precision solve ----------------------------
    X = None
    for _ in range(a.warmup):
        X, info = tmv_amd.gesv_mixed(A, b, tol=1e-12, max_iter=50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = []
    for _ in range(a.steps):
        X, info = tmv_amd.gesv_mixed(A, b, tol=1e-12, max_iter=50)
        infos.append(info)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / a.steps
    flops = 2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs
    value = flops / t_mixed / 1e12
    be_mixed = backward_error_gpu(A, X, b, torch)

    # ---- plain FP64 GPU LU solve on the same A, b ---------------------------
    t64 = []
    be64 = None
    for _ in range(a.fp64_reps):
        LU = A.clone()
        x64 = b.clone()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        piv, finfo = tmv_amd.getrf(LU)
        tmv_amd.getrs(LU, piv, x64)
        torch.cuda.synchronize()
        t64.append(time.perf_counter() - t1)
        be64 = backward_error_gpu(A, x64, b, torch)
        del LU
    t_fp64 = min(t64)
    del A
    torch.cuda.empty_cache()

    # ---- DGEMM 8192^3 (CCC, CRC, RCC) with HIP events -----------------------
    gn = a.gemm_n
    GA = torch.empty(gn * gn, dtype=torch.float64, device="cuda")
    GB = torch.empty(gn * gn, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(GA, SEED + 20)
    tmv_amd.fill_uniform(GB, SEED + 21)
    GC = tmv_amd.empty_colmajor(gn, gn, torch.float64)
    gemm = {}
    for lay in ("CC", "CR", "RC"):
        Av = GA.view(gn, gn).t() if lay[0] == "C" else GA.view(gn, gn)
        Bv = GB.view(gn, gn).t() if lay[1] == "C" else GB.view(gn, gn)
        tmv_amd.gemm(Av, Bv, GC)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(a.gemm_reps):
            tmv_amd.gemm(Av, Bv, GC)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / a.gemm_reps
        tf = 2.0 * gn ** 3 / ms / 1e9
        name = {"CC": "CCC", "CR": "CRC", "RC": "RCC"}[lay]
        gemm[name] = {"ms": round(ms, 3), "tflops": round(tf, 2),
                      "frac_peak": round(tf / PEAK_FP64_TFLOPS, 4)}
    del GA, GB, GC
    torch.cuda.empty_cache()

    cpu = None
    if not a.no_cpu_baseline:
        try:
            cpu = cpu_baseline()
        except Exception as ex:          # reported, never silently replaced
            cpu = {"value": None, "unit": "TFLOP/s", "cores": None, "kind": "reference",
                   "sample": f"failed: {ex}"}

    ccc = gemm["CCC"]
    out = {
        "metric": "LU-solve TFLOP/s: FP32-MFMA+IR vs FP64 at <=1e-12 backward err; DGEMM % peak",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factorisation + f64 refinement",
        "data": "synthetic: A = U diag(s) V^T (QR of seeded Gaussians), b ~ N(0,1)",
        "config": {"workload": "mixed-precision LU solve + iterative refinement",
                   "n": n, "nrhs": nrhs, "kappa": a.kappa, "target_backward_error": 1e-12,
                   "parallelism": "single GPU"},
        "mixed": {"backward_error": be_mixed,
                  "iterations": infos[-1]["iterations"],
                  "factor_ms": round(infos[-1]["factor_ms"], 3),
                  "refine_ms": round(infos[-1]["refine_ms"], 3)},
        "fp64_lu_solve": {"ms": round(t_fp64 * 1e3, 3),
                          "tflops": round(flops / t_fp64 / 1e12, 3),
                          "backward_error": be64},
        "speedup_vs_fp64": round(t_fp64 / t_mixed, 3),
        "dgemm_8192": gemm,
        "roofline": {"bound": "mfma", "achieved": ccc["tflops"], "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": ccc["frac_peak"],
                     "traffic": DGEMM_8192_CCC_TRAFFIC if gn == 8192 else None,
                     "kernel": "gemm_kernel f64 128x128x16 (DGEMM 8192^3, CCC)",
                     "flops_per_launch": 2 * gn ** 3},
        "cpu_baseline": cpu,
        "setup_s": round(setup_s, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
