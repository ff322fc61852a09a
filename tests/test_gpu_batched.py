"""Batched small-system LU + solve, stored-factor solves and batched GEMM on
the GPU through the C ABI (include/tmvx_batched.h), against the oracle
(TMV LU_Decompose / LU_LDivEq / MultMM restated) per system.

Bar: pivots identical to the oracle's on tie-free inputs (pivot gap above
64 n eps), ||P A - L U|| <= 8 n eps ||A||, backward error of every solution
column <= 8 n eps, GEMM within n eps |A||B| elementwise; every schedule
gives bit-identical results (same arithmetic, different thread layout);
a hipGraph replay gives the eager result every time."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS64, EPS32, lu_residual, backward_error, gemm_bound_ok
from tmv_amd import rng

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def make_batch(ns, nrhs, dtype, seed, shift_every=2):
    """A_i uniform[-1,1] (every shift_every-th also + n I, the batched
    workload's matrices), B_i uniform[-1,1]."""
    b = tmv_amd.Batch(ns, nrhs, dtype=dtype)
    As, Bs = [], []
    for i, (n, r) in enumerate(zip(b.ns, b.nrhs)):
        A = rng.uniform_matrix(seed + 2 * i, int(n), int(n))
        if shift_every and i % shift_every == 1:
            A = A + n * np.eye(n)
        B = rng.uniform_matrix(seed + 2 * i + 1, int(n), int(r)) if r else np.zeros((n, 0))
        As.append(A)
        Bs.append(B)
        b.a(i).copy_(torch.from_numpy(A).to(dtype))
        if r:
            b.b(i).copy_(torch.from_numpy(B).to(dtype))
    return b, As, Bs


EDGE_NS = [1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 32, 33, 47, 63, 64, 65, 96, 127, 128]


def mixed_sizes(count, seed):
    g = np.random.default_rng(seed)
    ns = np.concatenate([EDGE_NS, g.integers(1, 129, count - len(EDGE_NS))]).astype(np.int32)
    nrhs = g.integers(0, 5, len(ns)).astype(np.int32)
    nrhs[::7] = 8
    return ns, nrhs


def check_batch(orc, b, As, Bs, eps, factor=8.0):
    LU = b.A.cpu().numpy()
    P = b.P.cpu().numpy()
    X = b.B.cpu().numpy()
    info = b.info.cpu().numpy()
    for i, n in enumerate(b.ns):
        n = int(n)
        r = int(b.nrhs[i])
        A = As[i]
        oA, oP = int(b.offA[i]), int(b.offP[i])
        lu = LU[oA:oA + n * n].reshape(n, n).T
        p = P[oP:oP + n]
        assert info[i] == 0, (i, n)
        _, Po, gaps = orc.getrf(A, gaps=True)
        ok = gaps > 64 * n * eps
        k0 = n if ok.all() else int(np.argmin(ok))
        assert np.array_equal(p[:k0], Po[:k0]), (i, n)
        assert lu_residual(A, lu, p) <= factor * n * eps, (i, n)
        oB = int(b.offB[i])
        x = X[oB:oB + n * r].reshape(r, n).T
        for c in range(r):
            be = backward_error(A, x[:, c], Bs[i][:, c])
            assert be <= factor * n * eps, (i, n, c, be)


def test_batched_gesv_f64_mixed_sizes(orc, dev):
    ns, nrhs = mixed_sizes(240, 11)
    print(f"batched f64 count={len(ns)}", flush=True)
    b, As, Bs = make_batch(ns, nrhs, torch.float64, SEED + 1000)
    plan = tmv_amd.BatchPlan(b)
    tmv_amd.gesv_batched(plan, b)
    torch.cuda.synchronize()
    check_batch(orc, b, As, Bs, EPS64)


def test_batched_gesv_f32_mixed_sizes(orc, dev):
    ns, nrhs = mixed_sizes(160, 12)
    print(f"batched f32 count={len(ns)}", flush=True)
    b, As, Bs = make_batch(ns, nrhs, torch.float32, SEED + 2000)
    As32 = [A.astype(np.float32).astype(np.float64) for A in As]
    Bs32 = [B.astype(np.float32).astype(np.float64) for B in Bs]
    plan = tmv_amd.BatchPlan(b)
    tmv_amd.gesv_batched(plan, b)
    torch.cuda.synchronize()
    check_batch(orc, b, As32, Bs32, EPS32)


def test_batched_schedules_bit_identical(dev):
    """Every schedule, split and grid computes the same arithmetic in the
    same order per element: results must agree bit for bit."""
    ns, nrhs = mixed_sizes(120, 13)
    b, _, _ = make_batch(ns, nrhs, torch.float64, SEED + 3000)
    A0, B0 = b.A.clone(), b.B.clone()
    plan = tmv_amd.BatchPlan(b)
    ref = None
    for sched in range(tmv_amd.batch_schedule_count()):
        for split in (0, 16, 32, 64):
            for wpc in (1, 2):
                print(f"schedule {sched} split {split} wpc {wpc}", flush=True)
                plan.set(sched, split, wpc)
                b.A.copy_(A0)
                b.B.copy_(B0)
                tmv_amd.gesv_batched(plan, b)
                out = (b.A.clone(), b.B.clone(), b.P.clone(), b.info.clone())
                if ref is None:
                    ref = out
                else:
                    for u, v in zip(out, ref):
                        assert torch.equal(u, v), (sched, split, wpc)


def test_batched_zero_columns(orc, dev):
    """TMV's zero-column rule per system: a column whose candidates are all
    below DBL_MIN is zeroed from the diagonal, piv[j] = j, info = j + 1."""
    ns = [5, 6, 40, 70, 128]
    b = tmv_amd.Batch(ns, [0] * len(ns))
    As = []
    for i, n in enumerate(ns):
        A = rng.uniform_matrix(SEED + 4000 + i, n, n)
        A[:, n // 2] = 0.0
        if n > 6:
            A[n // 3:, n // 3] = 1e-310          # underflow candidates only
        As.append(A)
        b.a(i).copy_(torch.from_numpy(A))
    plan = tmv_amd.BatchPlan(b)
    tmv_amd.gesv_batched(plan, b)
    torch.cuda.synchronize()
    info = b.info.cpu().numpy()
    for i, n in enumerate(ns):
        LUo, Po, gaps = orc.getrf(As[i], gaps=True)
        lu = b.a(i).cpu().numpy()
        p = b.piv(i).cpu().numpy()
        first = int(np.nonzero(np.abs(np.diag(LUo)) == 0)[0][0]) + 1
        assert info[i] == first, (n, info[i], first)
        ok = gaps > 64 * n * EPS64
        k0 = n if ok.all() else int(np.argmin(ok))
        assert np.array_equal(p[:k0], Po[:k0])
        assert lu_residual(As[i], lu, p) <= 8 * n * EPS64


def test_batched_tuned_plan(dev):
    """The autotuner picks a schedule; its result equals the default's."""
    g = np.random.default_rng(14)
    ns = g.integers(8, 129, 2000).astype(np.int32)
    nrhs = g.integers(1, 5, 2000).astype(np.int32)
    print("tuned plan count=2000", flush=True)
    b, _, _ = make_batch(ns, nrhs, torch.float64, SEED + 5000, shift_every=1)
    A0, B0 = b.A.clone(), b.B.clone()
    p0 = tmv_amd.BatchPlan(b)
    tmv_amd.gesv_batched(p0, b)
    ref = (b.A.clone(), b.B.clone(), b.P.clone())
    pt = tmv_amd.BatchPlan(b, tune=True)
    assert "autotuned" in pt.describe()
    b.A.copy_(A0)
    b.B.copy_(B0)
    tmv_amd.gesv_batched(pt, b)
    for u, v in zip((b.A, b.B, b.P), ref):
        assert torch.equal(u, v)


def test_batched_graph_replays(dev):
    """One captured launch replayed with eager work in between: every replay
    equals the eager result."""
    g = np.random.default_rng(15)
    ns = g.integers(8, 129, 500).astype(np.int32)
    nrhs = g.integers(1, 5, 500).astype(np.int32)
    b, _, _ = make_batch(ns, nrhs, torch.float64, SEED + 6000, shift_every=1)
    A0, B0 = b.A.clone(), b.B.clone()
    plan = tmv_amd.BatchPlan(b)
    tmv_amd.gesv_batched(plan, b)
    torch.cuda.synchronize()
    ref = (b.A.clone(), b.B.clone(), b.P.clone(), b.info.clone())
    s = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    b.A.copy_(A0)
    b.B.copy_(B0)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        tmv_amd.gesv_batched(plan, b, stream=s)
    for rep in range(3):
        print(f"graph replay {rep}", flush=True)
        b.A.copy_(A0)
        b.B.copy_(B0)
        b.P.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for u, v in zip((b.A, b.B, b.P, b.info), ref):
            assert torch.equal(u, v), rep
        _ = (b.A * 2).sum().item()       # unrelated eager work between replays


def test_batched_rhs_beyond_wave_limit(orc, dev):
    """Small systems with more right-hand sides than a wave schedule holds
    run on the workgroup schedule."""
    ns = [4, 9, 20, 31]
    nrhs = [12, 16, 10, 16]
    b, As, Bs = make_batch(ns, nrhs, torch.float64, SEED + 7000)
    plan = tmv_amd.BatchPlan(b)
    tmv_amd.gesv_batched(plan, b)
    torch.cuda.synchronize()
    check_batch(orc, b, As, Bs, EPS64)


@pytest.mark.parametrize("trans", [False, True])
def test_getrs_batched(orc, dev, trans):
    ns = [1, 5, 33, 64, 100, 128, 200, 256]
    nrhs = [1, 2, 1, 3, 1, 4, 2, 1]
    b, As, Bs = make_batch(ns, nrhs, torch.float64, SEED + 8000 + int(trans))
    # factors: the batched factorisation for n <= 128, the blocked one above
    B0 = b.B.clone()
    for i, n in enumerate(ns):
        if n > 128:
            piv, info = tmv_amd.getrf(b.a(i))
            b.piv(i).copy_(piv)
    small = [i for i, n in enumerate(ns) if n <= 128]
    f = tmv_amd.Batch([ns[i] for i in small], [0] * len(small))
    for k, i in enumerate(small):
        f.a(k).copy_(b.a(i))
    tmv_amd.gesv_batched(tmv_amd.BatchPlan(f), f)
    for k, i in enumerate(small):
        b.a(i).copy_(f.a(k))
        b.piv(i).copy_(f.piv(k))
    b.B.copy_(B0)
    print(f"getrs_batched trans={trans}", flush=True)
    tmv_amd.getrs_batched(b, trans=trans)
    torch.cuda.synchronize()
    for i, n in enumerate(ns):
        x = b.b(i).cpu().numpy()
        A = As[i].T if trans else As[i]
        for c in range(nrhs[i]):
            assert backward_error(A, x[:, c], Bs[i][:, c]) <= 8 * n * EPS64, (i, n, c)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gemm_batched(dev, dtype):
    g = np.random.default_rng(16)
    cnt = 150
    m = g.integers(1, 129, cnt).astype(np.int32)
    n = g.integers(1, 129, cnt).astype(np.int32)
    k = g.integers(1, 129, cnt).astype(np.int32)
    m[:4], n[:4], k[:4] = [1, 16, 17, 128], [1, 16, 33, 128], [1, 4, 5, 128]
    As, Bs, Cs = [], [], []
    for i in range(cnt):
        As.append(torch.from_numpy(rng.uniform_matrix(SEED + 9000 + 3 * i, m[i], k[i])))
        Bs.append(torch.from_numpy(rng.uniform_matrix(SEED + 9001 + 3 * i, k[i], n[i])))
        Cs.append(torch.from_numpy(rng.uniform_matrix(SEED + 9002 + 3 * i, m[i], n[i])))
    dA = [tmv_amd.colmajor(a.to(dtype).cuda()) for a in As]
    dB = [tmv_amd.colmajor(x.to(dtype).cuda()) for x in Bs]
    dC = [tmv_amd.colmajor(c.to(dtype).cuda()) for c in Cs]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device=dev)
    dm, dn, dk = (torch.as_tensor(v, device=dev) for v in (m, n, k))
    eps = EPS64 if dtype == torch.float64 else EPS32
    print(f"gemm_batched {dtype}", flush=True)
    tmv_amd.gemm_batched(dm, dn, dk, ptr(dA), ptr(dB), ptr(dC), alpha=-0.5, beta=2.0,
                         dtype=dtype)
    torch.cuda.synchronize()
    for i in range(cnt):
        A = As[i].to(dtype).double().numpy()
        B = Bs[i].to(dtype).double().numpy()
        C0 = Cs[i].to(dtype).double().numpy()
        ref = -0.5 * (np.asarray(A, np.longdouble) @ np.asarray(B, np.longdouble)) + 2.0 * C0
        ok, worst = gemm_bound_ok(dC[i].double().cpu().numpy(), ref, A, B, int(k[i]), eps,
                                  factor=1.0, C0=C0, beta=2.0, alpha=-0.5)
        assert ok, (i, worst)
