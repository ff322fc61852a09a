"""GEMM tile-configuration sweep (the tuning tool).

    python -m tmv_amd.tune [--n 8192] [--dtype f64|f32|both] [--reps 3]
    python -m tmv_amd.tune --getrf 16384 [--dtype f32] [--reps 2] [--nopiv]
    python -m tmv_amd.tune --timeline RESULTS.db
    python -m tmv_amd.tune --kstats RESULTS.db
    python -m tmv_amd.tune --pmcstats RESULTS.db
    python -m tmv_amd.tune --trsv 16384 [--dtype f32] [--reps 3]
    python -m tmv_amd.tune --batched 20000 [--dtype f64] [--reps 3]

Times every configuration of the tile table (include/tmvx_tune.h) on each
operand layout with HIP events and prints TFLOP/s; --json writes the table.
--getrf times the blocked LU (uniform[-1,1] matrix of that order) instead,
for profiling its kernels (panel, interchanges, trailing updates); with
--nopiv the LU without interchanges of that matrix plus 2n I.
--timeline reads a rocprofv3 kernel-trace database (rocpd SQLite) of such
a run and prints, for the last LU in it (from its getrf_init_kernel on),
each stream's busy time per kernel family and its idle gaps.
"""
from __future__ import annotations

import argparse
import ctypes
import json

import torch

import tmv_amd

_i64, _int, _vp = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p


def _bind():
    L = tmv_amd.lib()
    L.tmvx_gemm_config_count.restype = _int
    L.tmvx_gemm_config_count.argtypes = [_int]
    L.tmvx_gemm_config_name.restype = ctypes.c_char_p
    L.tmvx_gemm_config_name.argtypes = [_int, _int]
    for suf, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
        f = getattr(L, "tmvx_gemm_cfg_" + suf)
        f.restype = _int
        f.argtypes = [_int, _int, _int, _i64, _i64, _i64, ct, _vp, _i64, _vp, _i64, ct, _vp,
                      _i64, _vp]
    return L


def sweep(n=8192, dtypes=("f64", "f32"), reps=3, layouts=("CC", "CR", "RC", "RR"),
          m=None, k=None, cfgs=None):
    L = _bind()
    m = m or n
    k = k or n
    out = []
    s = torch.cuda.current_stream().cuda_stream
    for suf in dtypes:
        dt = torch.float64 if suf == "f64" else torch.float32
        eb = 8 if suf == "f64" else 4
        A = torch.empty(m * k, dtype=torch.float64, device="cuda")
        B = torch.empty(k * n, dtype=torch.float64, device="cuda")
        tmv_amd.fill_uniform(A, 1)
        tmv_amd.fill_uniform(B, 2)
        A, B = A.to(dt), B.to(dt)
        C = torch.empty(m * n, dtype=dt, device="cuda")
        fn = getattr(L, "tmvx_gemm_cfg_" + suf)
        for cfg in (cfgs if cfgs is not None else range(L.tmvx_gemm_config_count(eb))):
            name = L.tmvx_gemm_config_name(eb, cfg).decode()
            for lay in layouts:
                la = 0 if lay[0] == "C" else 1
                lb = 0 if lay[1] == "C" else 1
                lda = m if la == 0 else k
                ldb = k if lb == 0 else n
                args = (cfg, la, lb, m, n, k, 1.0, A.data_ptr(), lda, B.data_ptr(), ldb, 0.0,
                        C.data_ptr(), m, s)
                rc = fn(*args)
                if rc != 0:
                    print(f"{suf} cfg {cfg} {lay}: rc={rc}")
                    continue
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    fn(*args)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                tf = 2.0 * m * n * k / ms / 1e9
                out.append({"dtype": suf, "cfg": cfg, "name": name, "layout": lay,
                            "m": m, "n": n, "k": k, "ms": ms, "tflops": tf})
                print(f"{suf} cfg {cfg} [{name}] {lay} {m}x{n}x{k}: {ms:8.3f} ms {tf:7.2f} TFLOP/s",
                      flush=True)
    return out


def _family(name: str) -> str:
    import re
    name = re.sub(r"\(.*", "", name)
    for k in ("swap_solve", "laswp_panel", "panel_kernel", "genp_diag", "genp_l21", "trsm_u12",
              "getrf_init", "trsv_inv", "trsv_persistent"):
        if k in name:
            return k
    m = re.search(r"GemmCfg<(\w+), (\d+), (\d+), (\d+)", name)
    if m:
        return f"gemm {m.group(1)} {m.group(2)}x{m.group(3)}x{m.group(4)}"
    return name[:60]


def timeline(db_path: str) -> None:
    """Per-stream busy time by kernel family for the last LU in a rocpd db."""
    import collections
    import sqlite3
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, stream_id, start, end from kernels order by start").fetchall()
    starts = [i for i, r in enumerate(rows) if "getrf_init_kernel" in r[0]]
    if not starts:
        raise SystemExit("no getrf_init_kernel dispatch in the database")
    sub = rows[starts[-1]:]
    t0, t1 = sub[0][2], max(r[3] for r in sub)
    print(f"LU window {(t1 - t0) / 1e6:.2f} ms, {len(sub)} dispatches")
    by_stream = collections.defaultdict(list)
    for r in sub:
        by_stream[r[1]].append(r)
    for sid, rs in sorted(by_stream.items()):
        busy, cnt = collections.Counter(), collections.Counter()
        for r in rs:
            busy[_family(r[0])] += (r[3] - r[2]) / 1e6
            cnt[_family(r[0])] += 1
        gaps = sum(max(0, rs[i + 1][2] - rs[i][3]) for i in range(len(rs) - 1)) / 1e6
        print(f"stream {sid}: {len(rs)} kernels, busy {sum(busy.values()):.2f} ms, "
              f"gaps {gaps:.2f} ms, span {(rs[-1][3] - rs[0][2]) / 1e6:.2f} ms")
        for k, v in busy.most_common():
            print(f"    {v:8.2f} ms  {cnt[k]:5d} x {1e3 * v / cnt[k]:8.1f} us  {k}")


def kstats(db_path: str, match: str = "tmvx") -> None:
    """Per-kernel dispatch statistics (durations end - start) of a rocpd db,
    kernels whose name contains `match`, by total time."""
    import collections
    import sqlite3
    db = sqlite3.connect(db_path)
    d = collections.defaultdict(list)
    for name, s0, s1 in db.execute("select name, start, end from kernels"):
        if match in name:
            d[name.split("(")[0].replace("void ", "")].append((s1 - s0) / 1e3)
    print("| kernel | calls | total ms | avg us | min us | max us |")
    print("|---|---|---|---|---|---|")
    for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{k}` | {len(v)} | {sum(v) / 1e3:.2f} | {sum(v) / len(v):.1f} | "
              f"{min(v):.1f} | {max(v):.1f} |")


def pmcstats(db_path: str, match: str = "tmvx") -> None:
    """Counter values per dispatch (summed over the device's blocks) of a
    rocprofv3 --pmc db, kernels whose name contains `match`: the mean over
    each kernel's dispatches."""
    import collections
    import sqlite3
    db = sqlite3.connect(db_path)
    q = ("select dispatch_id, kernel_name, counter_name, sum(value), max(duration) "
         "from counters_collection group by dispatch_id, counter_name")
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    durs = collections.defaultdict(dict)
    for did, kn, cn, v, du in db.execute(q):
        if match not in kn:
            continue
        k = kn.split("(")[0].replace("void ", "")
        per[k][cn].append(v)
        durs[k][did] = du
    for k, cs in per.items():
        d = list(durs[k].values())
        print(f"{k}: {len(d)} dispatches, mean duration {sum(d) / len(d) / 1e3:.1f} us")
        for cn in sorted(cs):
            v = cs[cn]
            print(f"    {cn:24s} {sum(v) / len(v):16.0f}")


def trsv_times(n: int, dtypes, reps: int) -> None:
    """Wall time of the public triangular solve (substitution) and the
    correction solve (inverted diagonal blocks) on the L and U factors of a
    uniform matrix's LU, one right-hand side."""
    import time
    A0 = torch.empty(n * n, dtype=torch.float64, device="cuda")
    tmv_amd.fill_uniform(A0, 5)
    for suf in dtypes:
        dt = torch.float64 if suf == "f64" else torch.float32
        LU = A0.to(dt).view(n, n).t()
        _, info = tmv_amd.getrf(LU)
        for name, fn in (("substitution", tmv_amd.trsm), ("block inverse", tmv_amd.trsm_blockinv)):
            for upper in (0, 1):
                for r in range(reps):
                    b = torch.ones(n, dtype=dt, device="cuda")
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    fn(LU, b, upper=upper, unit=not upper)
                    torch.cuda.synchronize()
                    print(f"trsv {suf} n={n} {name} {'U' if upper else 'L'} rep {r}: "
                          f"{1e3 * (time.perf_counter() - t0):.3f} ms", flush=True)


def batched_runs(count: int, dtypes, reps: int) -> None:
    """The batched factor + solve of `count` systems, n ~ U{8..128}, nrhs ~
    U{1..4}, uniform entries + n on the diagonal (the benchmark's mix), with
    the default plan, `reps` launches per precision (for counter passes)."""
    import time
    import numpy as np
    g = np.random.default_rng(0x7B)
    ns = g.integers(8, 129, count).astype(np.int32)
    nrhs = g.integers(1, 5, count).astype(np.int32)
    for suf in dtypes:
        dt = torch.float64 if suf == "f64" else torch.float32
        b = tmv_amd.Batch(ns, nrhs, dtype=dt)
        tmp = torch.empty(b.A.numel(), dtype=torch.float64, device="cuda")
        tmv_amd.fill_uniform(tmp, 0x7C)
        A0 = tmp.to(dt)
        diag = np.concatenate([b.offA[i] + np.arange(n) * (n + 1) for i, n in enumerate(ns)])
        A0[torch.as_tensor(diag, device="cuda")] += torch.as_tensor(
            np.repeat(ns.astype(np.float64), ns), device="cuda").to(dt)
        tmp = torch.empty(b.B.numel(), dtype=torch.float64, device="cuda")
        tmv_amd.fill_uniform(tmp, 0x7D)
        B0 = tmp.to(dt)
        plan = tmv_amd.BatchPlan(b)
        for r in range(reps):
            b.A.copy_(A0)
            b.B.copy_(B0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tmv_amd.gesv_batched(plan, b)
            torch.cuda.synchronize()
            print(f"batched {suf} {count} systems rep {r}: "
                  f"{1e3 * (time.perf_counter() - t0):.3f} ms ({plan.describe()})", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--k", type=int, default=None)
    ap.add_argument("--dtype", default="both")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--layouts", default="CC,CR,RC,RR")
    ap.add_argument("--json", default=None)
    ap.add_argument("--cfg", type=int, action="append", default=None,
                    help="only these configuration ids")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="instead: time a plain device-to-device copy of this size "
                         "(counter calibration)")
    ap.add_argument("--getrf", type=int, default=0,
                    help="instead: time the LU of a uniform matrix of this order")
    ap.add_argument("--nopiv", action="store_true",
                    help="with --getrf: the LU without interchanges (of A + 2n I)")
    ap.add_argument("--timeline", default=None,
                    help="instead: summarise the last LU of a rocprofv3 kernel-trace db")
    ap.add_argument("--kstats", default=None,
                    help="instead: per-kernel statistics of a rocprofv3 kernel-trace db")
    ap.add_argument("--pmcstats", default=None,
                    help="instead: per-kernel counter means of a rocprofv3 --pmc db")
    ap.add_argument("--trsv", type=int, default=0,
                    help="instead: time the triangular solves on an LU of this order")
    ap.add_argument("--batched", type=int, default=0,
                    help="instead: run the batched factor + solve on this many systems")
    a = ap.parse_args()
    if a.timeline:
        timeline(a.timeline)
        return
    if a.kstats:
        kstats(a.kstats)
        return
    if a.pmcstats:
        pmcstats(a.pmcstats)
        return
    if a.trsv:
        trsv_times(a.trsv, ("f64", "f32") if a.dtype == "both" else (a.dtype,), a.reps)
        return
    if a.batched:
        batched_runs(a.batched, ("f64", "f32") if a.dtype == "both" else (a.dtype,), a.reps)
        return
    if a.getrf:
        import time
        n = a.getrf
        A0 = torch.empty(n * n, dtype=torch.float64, device="cuda")
        tmv_amd.fill_uniform(A0, 3)
        for suf in (("f64", "f32") if a.dtype == "both" else (a.dtype,)):
            dt = torch.float64 if suf == "f64" else torch.float32
            A = A0.to(dt).view(n, n).t()
            if a.nopiv:
                idx = torch.arange(n, device="cuda")
                A[idx, idx] += 2 * n
            for r in range(a.reps):
                M = A.clone()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _, info = (tmv_amd.getrf_nopiv if a.nopiv else tmv_amd.getrf)(M)
                torch.cuda.synchronize()
                print(f"getrf {suf} n={n} rep {r}: {1e3 * (time.perf_counter() - t0):.2f} ms "
                      f"info={info}", flush=True)
                del M
        return
    if a.copy_mib:
        nb = a.copy_mib * (1 << 20)
        src = torch.empty(nb // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(a.reps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {a.copy_mib} MiB x {a.reps}")
        return
    dts = ("f64", "f32") if a.dtype == "both" else (a.dtype,)
    res = sweep(a.n, dts, a.reps, tuple(a.layouts.split(",")), a.m, a.k, a.cfg)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
