"""Adjoint of the LU solve (include/tmvx.h tmvx_gesv_adjoint_*,
include/tmvx_batched.h tmvx_gesv_adjoint_batched_*), through the C ABI.

The autograd wrapper below exists in the tests only: forward = batched
factor + solve, backward = the hand-written adjoint kernel (stored factors,
or recomputed ones).  Checked against
  - the gradient from TMV's own transposed solve: grad_b = g % A (TMV
    LU_RDivEq through the real TMV build when present, else the oracle),
    grad_A = -grad_b x^T;
  - FP64 central finite differences of loss = sum(w .* x), h = 1e-6, with an
    independent FP64 solver (numpy);
  - bit-identical results with and without factor recomputation.
Matrices of the differentiation workload: A = Q diag(s) Q^T + n I, Q from
the QR of a seeded Gaussian, s ~ U[0.1, 10]; b, w ~ N(0, 1)."""
import numpy as np
import pytest
import torch

import tmv_amd
from checks import EPS64, EPS32

pytestmark = pytest.mark.gpu
SEED = 0x7A15EED


def diff_matrix(n, seed):
    g = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(g.standard_normal((n, n)))
    s = g.uniform(0.1, 10.0, n)
    return (Q * s) @ Q.T + n * np.eye(n)


class BatchedSolve(torch.autograd.Function):
    """x_i = A_i^-1 b_i over a packed batch; backward by the adjoint kernel."""

    @staticmethod
    def forward(ctx, A_flat, B_flat, ns, nrhs, recompute):
        fact = tmv_amd.Batch(ns, nrhs, dtype=A_flat.dtype)
        fact.A.copy_(A_flat)
        fact.B.copy_(B_flat)
        info = tmv_amd.getrf_batched(fact)
        assert int(info.abs().sum()) == 0
        tmv_amd.getrs_batched(fact)
        X = fact.B.clone()
        ctx.ns, ctx.nrhs, ctx.recompute = ns, nrhs, recompute
        if recompute:                     # keep A, drop the factors
            ctx.fact = tmv_amd.Batch(ns, nrhs, dtype=A_flat.dtype, A=A_flat.detach().clone())
        else:
            ctx.fact = fact
        ctx.save_for_backward(X)
        return X

    @staticmethod
    def backward(ctx, gX):
        (X,) = ctx.saved_tensors
        dt = X.dtype
        Xb = tmv_amd.Batch(ctx.ns, ctx.nrhs, dtype=dt, B=X.contiguous())
        Gb = tmv_amd.Batch(ctx.ns, ctx.nrhs, dtype=dt, B=gX.contiguous())
        gB = tmv_amd.Batch(ctx.ns, ctx.nrhs, dtype=dt)
        gA = tmv_amd.Batch(ctx.ns, ctx.nrhs, dtype=dt)
        tmv_amd.gesv_adjoint_batched(ctx.fact, Xb, Gb, gB, gA, recompute=ctx.recompute)
        return gA.A, gB.B, None, None, None


def make(ns, nrhs, seed, dtype=torch.float64):
    lay = tmv_amd.Batch(ns, nrhs, dtype=dtype)
    As, Bs, Ws = [], [], []
    g = np.random.default_rng(seed + 1)
    for i, (n, r) in enumerate(zip(ns, nrhs)):
        A = diff_matrix(n, seed + 10 * i)
        B = g.standard_normal((n, r))
        W = g.standard_normal((n, r))
        As.append(A); Bs.append(B); Ws.append(W)
        lay.a(i).copy_(torch.from_numpy(A).to(dtype))
        lay.b(i).copy_(torch.from_numpy(B).to(dtype))
    Wflat = torch.zeros_like(lay.B)
    for i, (n, r) in enumerate(zip(ns, nrhs)):
        o = int(lay.offB[i])
        Wflat[o:o + n * r] = torch.from_numpy(Ws[i].T.reshape(-1)).to(dtype)
    return lay, As, Bs, Ws, Wflat


def run(lay, Wflat, ns, nrhs, recompute):
    A = lay.A.clone().requires_grad_(True)
    B = lay.B.clone().requires_grad_(True)
    X = BatchedSolve.apply(A, B, ns, nrhs, recompute)
    loss = (X * Wflat).sum()
    loss.backward()
    torch.cuda.synchronize()
    return X.detach(), A.grad, B.grad


NS = [1, 7, 64, 100, 128, 129, 200, 256]
NRHS = [1, 2, 1, 5, 3, 1, 4, 1]


def blocks(lay, flatA, flatB, i):
    n, r = int(lay.ns[i]), int(lay.nrhs[i])
    oA, oB = int(lay.offA[i]), int(lay.offB[i])
    gA = flatA[oA:oA + n * n].view(n, n).t().double().cpu().numpy()
    gB = flatB[oB:oB + n * r].view(r, n).t().double().cpu().numpy()
    return gA, gB


def test_adjoint_vs_tmv_transposed_solve(orc, dev):
    print("adjoint vs TMV transposed solve", flush=True)
    lay, As, Bs, Ws, Wflat = make(NS, NRHS, SEED + 100)
    X, gAflat, gBflat = run(lay, Wflat, NS, NRHS, recompute=False)
    ref = orc.ref_lib()
    for i, (n, r) in enumerate(zip(NS, NRHS)):
        gA, gB = blocks(lay, gAflat, gBflat, i)
        LU, P = orc.getrf(As[i], lib=ref)[:2]
        gb_ref = orc.getrs(LU, P, Ws[i], trans=True, lib=ref)        # w % A, per column
        x_ref = orc.getrs(LU, P, Bs[i], lib=ref)
        gA_ref = -gb_ref @ x_ref.T
        scale = np.abs(gb_ref).max()
        assert np.abs(gB - gb_ref).max() <= 64 * n * EPS64 * scale, (i, n)
        assert np.abs(gA - gA_ref).max() <= 64 * n * EPS64 * np.abs(gA_ref).max(), (i, n)


def test_adjoint_finite_differences(dev):
    print("adjoint vs central finite differences", flush=True)
    ns, nrhs = [5, 64, 256], [2, 1, 1]
    lay, As, Bs, Ws, Wflat = make(ns, nrhs, SEED + 200)
    X, gAflat, gBflat = run(lay, Wflat, ns, nrhs, recompute=False)
    h = 1e-6
    g = np.random.default_rng(3)
    for i, n in enumerate(ns):
        gA, gB = blocks(lay, gAflat, gBflat, i)
        loss = lambda A, B: float((np.linalg.solve(A, B) * Ws[i]).sum())
        for _ in range(4):
            j, k = g.integers(0, n, 2)
            Ap, Am = As[i].copy(), As[i].copy()
            Ap[j, k] += h; Am[j, k] -= h
            fd = (loss(Ap, Bs[i]) - loss(Am, Bs[i])) / (2 * h)
            assert abs(fd - gA[j, k]) <= 1e-7 + 1e-5 * abs(fd), (n, j, k, fd, gA[j, k])
            c = g.integers(0, Bs[i].shape[1])
            Bp, Bm = Bs[i].copy(), Bs[i].copy()
            Bp[j, c] += h; Bm[j, c] -= h
            fd = (loss(As[i], Bp) - loss(As[i], Bm)) / (2 * h)
            assert abs(fd - gB[j, c]) <= 1e-7 + 1e-5 * abs(fd), (n, j, c, fd, gB[j, c])


def test_adjoint_recompute_bit_identical(dev):
    print("adjoint stored vs recomputed factors", flush=True)
    lay, As, Bs, Ws, Wflat = make(NS, NRHS, SEED + 300)
    _, gA1, gB1 = run(lay, Wflat, NS, NRHS, recompute=False)
    _, gA2, gB2 = run(lay, Wflat, NS, NRHS, recompute=True)
    assert torch.equal(gA1, gA2) and torch.equal(gB1, gB2)


def test_adjoint_f32(dev):
    print("adjoint float32", flush=True)
    ns, nrhs = [16, 128, 256], [1, 2, 1]
    lay64, As, Bs, Ws, W64 = make(ns, nrhs, SEED + 400)
    _, gA64, gB64 = run(lay64, W64, ns, nrhs, recompute=False)
    lay32, _, _, _, W32 = make(ns, nrhs, SEED + 400, dtype=torch.float32)
    _, gA32, gB32 = run(lay32, W32, ns, nrhs, recompute=False)
    for a, b in ((gA32, gA64), (gB32, gB64)):
        err = (a.double() - b).abs().max() / b.abs().max()
        assert err <= 64 * 256 * EPS32, float(err)


@pytest.mark.parametrize("ns,nrhs", [([30, 130], [2, 1]), ([64, 200], [9, 6])])
def test_adjoint_accumulate(dev, ns, nrhs):
    """accumulate = 1 adds to gradA (right-hand sides in blocks of four:
    nrhs 9 and 6 take several blocks)."""
    lay, As, Bs, Ws, Wflat = make(ns, nrhs, SEED + 500)
    fact = tmv_amd.Batch(ns, nrhs)
    fact.A.copy_(lay.A)
    fact.B.copy_(lay.B)
    tmv_amd.getrf_batched(fact)
    tmv_amd.getrs_batched(fact)
    Xb = tmv_amd.Batch(ns, nrhs, B=fact.B.clone())
    Gb = tmv_amd.Batch(ns, nrhs, B=Wflat.clone())
    gB = tmv_amd.Batch(ns, nrhs)
    gA = tmv_amd.Batch(ns, nrhs)
    tmv_amd.gesv_adjoint_batched(fact, Xb, Gb, gB, gA)
    first = gA.A.clone()
    gA.A.add_(1.0)
    tmv_amd.gesv_adjoint_batched(fact, Xb, Gb, gB, gA, accumulate=True)
    torch.cuda.synchronize()
    assert torch.allclose(gA.A, 2 * first + 1.0, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("n,nrhs", [(1000, 1), (700, 6)])
def test_adjoint_single_large(dev, n, nrhs):
    """One large system: transposed solve + rank-nrhs GEMM update."""
    print(f"single adjoint n={n} nrhs={nrhs}", flush=True)
    g = np.random.default_rng(n)
    A = g.uniform(-1, 1, (n, n)) + n ** 0.5 * np.eye(n)
    Bm = g.standard_normal((n, nrhs))
    W = g.standard_normal((n, nrhs))
    LU = tmv_amd.colmajor(torch.from_numpy(A).cuda())
    piv, info = tmv_amd.getrf(LU)
    X = tmv_amd.colmajor(torch.from_numpy(Bm).cuda())
    tmv_amd.getrs(LU, piv, X)
    G = tmv_amd.colmajor(torch.from_numpy(W).cuda())
    gB, gA = tmv_amd.gesv_adjoint(LU, piv, X, G)
    torch.cuda.synchronize()
    gb_ref = np.linalg.solve(A.T, W)
    x_ref = np.linalg.solve(A, Bm)
    ga_ref = -gb_ref @ x_ref.T
    assert np.abs(gB.cpu().numpy() - gb_ref).max() <= 1e-11 * np.abs(gb_ref).max()
    assert np.abs(gA.cpu().numpy() - ga_ref).max() <= 1e-11 * np.abs(ga_ref).max()
