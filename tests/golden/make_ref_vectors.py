"""Write tests/golden/tmv_ref_vectors.npz: outputs of the real TMV 0.76
library (built from the reference sources into oracle/_ref by
oracle/Makefile.ref) on seeded inputs.

    python tests/golden/make_ref_vectors.py

Inputs come from tmv_amd.rng (SplitMix64), so they can be regenerated
anywhere; only TMV's outputs are stored.  Large outputs are stored as
SHA-256 digests of their column-major float64 bytes.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import oracle  # noqa: E402
from tmv_amd import rng  # noqa: E402
import fixtures as F  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "tmv_ref_vectors.npz")
SEED = 0x7A15EED


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def digest_c(z: np.ndarray) -> str:
    """SHA-256 of a complex matrix's column-major interleaved bytes."""
    return hashlib.sha256(np.asfortranarray(z, np.complex128).tobytes(order="F")).hexdigest()


# (tag, conj_a, conj_b, alpha, add) for the complex GEMM vectors
ZGEMM_VARIANTS = [("n_n_real", 0, 0, 1.0, False), ("c_n_add", 1, 0, -0.5, True),
                  ("n_c_cplx", 0, 1, 0.3 + 0.7j, False), ("c_c_add", 1, 1, 2.0, True)]


HADAMARD_SIZES = (64, 256, 1024)
RECT_SHAPES = ((5, 3), (3, 5), (100, 37), (37, 100), (300, 130), (130, 300), (1, 7), (7, 1))


def hadamard_rows(n: int) -> np.ndarray:
    """Sylvester-Hadamard matrix of order n (a power of 2) with its rows in
    the order argsort(u), u uniform from the documented generator."""
    H = np.ones((1, 1))
    while H.shape[0] < n:
        H = np.block([[H, H], [H, -H]])
    return H[np.argsort(rng.uniform(SEED + 600 + n, n), kind="stable")]


def cfg2_matrix() -> np.ndarray:
    """The n = 8192 uniform [-1,1] matrix of the plain FP64 configuration at
    scale: tmvx_fill_uniform_f64(seed + 50) over n*n values read column-major."""
    n = 8192
    return rng.uniform(SEED + 50, n * n, -1.0, 1.0).reshape(n, n).T.copy()


def main() -> None:
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libtmv_ref.so is not built")
    v = {}
    # GEMM, all four operand layouts (A,B given column-major; the layout
    # flag says how TMV views the same values).
    A = rng.uniform_matrix(SEED + 101, 64, 80)
    B = rng.uniform_matrix(SEED + 102, 80, 48)
    for ar in (0, 1):
        for br in (0, 1):
            v[f"gemm_64x48x80_{ar}{br}"] = oracle.gemm(A, B, a_rowmajor=ar, b_rowmajor=br, lib=ref)
    C0 = rng.uniform_matrix(SEED + 103, 64, 48)
    v["gemm_64x48x80_add"] = oracle.gemm(A, B, C=C0, alpha=-0.5, add=True, lib=ref)
    # Complex GEMM (TMV's four-real-product block path: aligned shape) as
    # digests, and one ragged (small-kernel path) output.
    A = rng.uniform_matrix(SEED + 111, 128, 160) + 1j * rng.uniform_matrix(SEED + 112, 128, 160)
    B = rng.uniform_matrix(SEED + 113, 160, 96) + 1j * rng.uniform_matrix(SEED + 114, 160, 96)
    C0 = rng.uniform_matrix(SEED + 115, 128, 96) + 1j * rng.uniform_matrix(SEED + 116, 128, 96)
    for tag, ca, cb, alpha, add in ZGEMM_VARIANTS:
        Z = oracle.ref_gemm_complex(A, B, C=C0, alpha=alpha, add=add, conj_a=ca, conj_b=cb)
        v[f"zgemm_128x96x160_{tag}_sha256"] = np.array(digest_c(Z))
    A = rng.uniform_matrix(SEED + 121, 37, 41) + 1j * rng.uniform_matrix(SEED + 122, 37, 41)
    B = rng.uniform_matrix(SEED + 123, 41, 29) + 1j * rng.uniform_matrix(SEED + 124, 41, 29)
    v["zgemm_37x29x41_conjA"] = oracle.ref_gemm_complex(A, B, alpha=0.3 + 0.7j, conj_a=True)
    # Complex LU pivots (TMV_ABS = modulus, first max on ties).
    for n in (5, 37, 100, 257):
        Z = rng.uniform_matrix(SEED + 230 + n, n, n) + 1j * rng.uniform_matrix(SEED + 240 + n, n, n)
        _, P = oracle.getrf_complex(Z, lib=ref)
        v[f"zlu_{n}_P"] = P
    # QR (QR_Decompose's beta and R) and QR division, tall and square.
    for (m, n) in ((37, 37), (100, 64)):
        Q = rng.uniform_matrix(SEED + 260 + m + n, m, n)
        QR, beta = oracle.ref_geqrf(Q)
        v[f"qr_{m}x{n}_beta"] = beta
        v[f"qr_{m}x{n}_R"] = np.triu(QR[:n])
        Bq = rng.uniform_matrix(SEED + 270 + m, m, 2)
        v[f"qr_{m}x{n}_x"] = oracle.ref_qr_solve(Q, Bq)
    # LU on uniform [-1,1] inputs without a diagonal shift (pivoting active).
    for n in (1, 2, 3, 5, 37, 64, 100, 130):
        M = rng.uniform_matrix(SEED + 200 + n, n, n)
        LU, P = oracle.getrf(M, lib=ref)
        v[f"lu_{n}_LU"] = LU
        v[f"lu_{n}_P"] = P
        b = rng.uniform(SEED + 300 + n, n, -1.0, 1.0)
        v[f"solve_{n}_x"] = oracle.ref_gesv(M, b)
        # m.inverse() and m.logDet(&sign) with LU division.
        v[f"inv_{n}"] = oracle.ref_inverse(M)
        v[f"logdet_{n}"] = np.array(oracle.ref_logdet(M))
    # LU_Decompose on non-square views (R = min(m, n) pivots): tall and wide.
    for (m, n) in RECT_SHAPES:
        M = rng.uniform_matrix(SEED + 700 + m + 3 * n, m, n)
        LU, P = oracle.getrf(M, lib=ref)
        v[f"lurect_{m}x{n}_P"] = P
        v[f"lurect_{m}x{n}_LU"] = LU
    # TMV's own fixtures (tests/TMV_TestMatrixDecomp.cpp): pivots + digest.
    for t in (0, 3, 4, 5, 6):
        LU, P = oracle.getrf(F.decomp_matrix(t), lib=ref)
        v[f"decomp_{t}_P"] = P
        v[f"decomp_{t}_LU"] = LU
    LU, P = oracle.getrf(F.big_div_100(), lib=ref)
    v["bigdiv_P"] = P
    v["bigdiv_LU"] = LU
    # The plain CPU-runnable case: 512^3 GEMM and an LU solve on A + 512 I.
    n = 512
    A = rng.uniform_matrix(SEED + 0, n, n)
    B = rng.uniform_matrix(SEED + 1, n, n)
    C = oracle.gemm(A, B, lib=ref)
    v["cfg1_gemm_sha256"] = np.array(digest(C))
    v["cfg1_gemm_colsum"] = C.sum(axis=0)
    As = A + n * np.eye(n)
    LU, P = oracle.getrf(As, lib=ref)
    v["cfg1_lu_P"] = P
    v["cfg1_lu_sha256"] = np.array(digest(LU))
    b = rng.uniform(SEED + 2, n, -1.0, 1.0)
    v["cfg1_x"] = oracle.ref_gesv(As, b)
    v["cfg1_logdet"] = np.array(oracle.ref_logdet(As))
    # Exact ties: Sylvester-Hadamard matrices with seeded row orders.  Every
    # elimination step is an exact tie among +-|pivot| candidates and all
    # arithmetic is exact, so TMV's first-index rule alone fixes the pivots
    # and the factors are exact (digest of the column-major bytes).
    for n in HADAMARD_SIZES:
        H = hadamard_rows(n)
        LU, P = oracle.getrf(H, lib=ref)
        v[f"hadamard_{n}_P"] = P
        v[f"hadamard_{n}_LU_sha256"] = np.array(digest(LU))
    # The plain FP64 configuration at scale: LU of the n = 8192 uniform
    # [-1,1] matrix (no shift) -- TMV's whole pivot sequence, and the
    # smallest relative gap between the two largest candidates over the 8192
    # steps (from the restatement, whose pivots equal TMV's).
    n = 8192
    A = cfg2_matrix()
    _, P = oracle.getrf(A, lib=ref)
    _, Po, gaps = oracle.getrf(A, gaps=True)
    assert np.array_equal(P, Po)
    v["cfg2_lu_P"] = P
    v["cfg2_lu_min_gap"] = np.array(gaps.min())
    np.savez_compressed(OUT, **v)
    print(f"wrote {OUT}: {len(v)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
