"""Pins the oracle (the CPU restatement in oracle/tmv_oracle.cpp) against
TMV's own known answers, against the real TMV library's outputs
(tests/golden/tmv_ref_vectors.npz, and oracle/_ref live where built), and
against what the mathematics fixes.  CPU only."""
import hashlib

import numpy as np
import pytest
import scipy.linalg as sla

import fixtures as F
from checks import (EPS64, EPS32, gemm_bound_ok, lu_residual, backward_error,
                    inverse_residual, inverse_close)
from tmv_amd import rng

EPS = 10 * EPS64          # TMV's test EPS (tests/TMV_Test.h:24)
SEED = 0x7A15EED


# ---------------------------------------------------------------- GEMM ----

def test_gemm_matrix_example_exact(orc):
    """examples/Matrix.cpp:297-303, 329-334: integer products, exact."""
    m1, m5 = F.matrix_example()
    K = F.KNOWN["matrix_example"]
    assert np.array_equal(orc.gemm(m1, m5), np.array(K["m1m5"], float))
    assert np.array_equal(orc.gemm(m5, m5), np.array(K["m5m5"], float))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (3, 5, 2), (16, 16, 16), (37, 21, 50),
                                   (64, 64, 64), (77, 77, 77), (130, 65, 97)])
@pytest.mark.parametrize("ar,br", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_vs_long_double(orc, m, n, k, ar, br):
    A = rng.uniform_matrix(SEED + 11, m, k)
    B = rng.uniform_matrix(SEED + 12, k, n)
    C = orc.gemm(A, B, a_rowmajor=ar, b_rowmajor=br)
    Cl = orc.gemm(A, B, a_rowmajor=ar, b_rowmajor=br, dtype=np.longdouble)
    ok, worst = gemm_bound_ok(C, Cl, A, B, k, EPS64, factor=1.0)
    assert ok, worst


def test_gemm_integer_inputs_bit_exact(orc):
    """Integer entries whose partial sums stay below 2^53 give exact
    products in any summation order."""
    A = np.floor(rng.uniform_matrix(SEED + 13, 100, 90, -50, 50))
    B = np.floor(rng.uniform_matrix(SEED + 14, 90, 70, -50, 50))
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    for ar in (0, 1):
        for br in (0, 1):
            assert np.array_equal(orc.gemm(A, B, a_rowmajor=ar, b_rowmajor=br), exact)


def test_gemm_add_and_alpha(orc):
    A = rng.uniform_matrix(SEED + 15, 40, 33)
    B = rng.uniform_matrix(SEED + 16, 33, 29)
    C0 = rng.uniform_matrix(SEED + 17, 40, 29)
    C = orc.gemm(A, B, C=C0, alpha=-0.5, add=True)
    Cl = C0.astype(np.longdouble) - 0.5 * (A.astype(np.longdouble) @ B.astype(np.longdouble))
    ok, worst = gemm_bound_ok(C, Cl, A, B, 33, EPS64, factor=2.0, C0=C0, beta=1.0, alpha=0.5)
    assert ok, worst
    # alpha == 0 and !add zeroes C (src/TMV_MultMM.cpp:429-431)
    assert np.array_equal(orc.gemm(A, B, C=C0, alpha=0.0), np.zeros_like(C0))


def test_gemm_planted_transpose_is_caught(orc):
    """The bound check is sharp enough to reject a transposed operand."""
    A = rng.uniform_matrix(SEED + 18, 32, 32)
    B = rng.uniform_matrix(SEED + 19, 32, 32)
    wrong = orc.gemm(A.T.copy(), B)
    ok, _ = gemm_bound_ok(wrong, A @ B, A, B, 32, EPS64, factor=2.0)
    assert not ok


def test_gemm_matches_tmv_ref_vectors(orc, ref_vectors):
    """Bit-for-bit with the real TMV on tile-aligned shapes (16x16 tiles,
    K blocks of 32: the restatement follows sse2_multmm_16_16_32's order)."""
    A = rng.uniform_matrix(SEED + 101, 64, 80)
    B = rng.uniform_matrix(SEED + 102, 80, 48)
    for ar in (0, 1):
        for br in (0, 1):
            C = orc.gemm(A, B, a_rowmajor=ar, b_rowmajor=br)
            assert np.array_equal(C, ref_vectors[f"gemm_64x48x80_{ar}{br}"])
    C0 = rng.uniform_matrix(SEED + 103, 64, 48)
    assert np.array_equal(orc.gemm(A, B, C=C0, alpha=-0.5, add=True),
                          ref_vectors["gemm_64x48x80_add"])


def test_gemm_cfg1_512_matches_tmv(orc, ref_vectors):
    """512^3 (the CPU-runnable configuration): identical bytes to TMV."""
    n = 512
    A = rng.uniform_matrix(SEED + 0, n, n)
    B = rng.uniform_matrix(SEED + 1, n, n)
    C = orc.gemm(A, B)
    h = hashlib.sha256(np.asfortranarray(C).tobytes(order="F")).hexdigest()
    assert h == str(ref_vectors["cfg1_gemm_sha256"])


# ------------------------------------------------------------------ LU ----

def test_det_fixtures(orc):
    """tests/TMV_TestMatrixDet.cpp: exact determinants within TMV's own
    tolerance k*NormSq(m)*EPS."""
    mats = F.det_matrices()
    for name, (true, k) in F.KNOWN["dets"].items():
        m = mats[name]
        LU, P = orc.getrf(m)
        d = orc.det_from_lu(LU, P)
        norm_m = mats["m11"] if name == "m12" else m
        tol = k * float(np.sum(norm_m ** 2)) * EPS
        assert abs(d - true) <= tol, (name, d, true)


def test_square_div_det_28800(orc):
    m, b = F.square_div_4x4()
    LU, P = orc.getrf(m)
    assert abs(orc.det_from_lu(LU, P) - F.KNOWN["square_div_det"]) < 1e-9 * 28800
    x = orc.getrs(LU, P, b)
    assert backward_error(m, x, b) < 8 * 4 * EPS64


# ------------------------------------------------------------- complex ----

ZGEMM_VARIANTS = [("n_n_real", 0, 0, 1.0, False), ("c_n_add", 1, 0, -0.5, True),
                  ("n_c_cplx", 0, 1, 0.3 + 0.7j, False), ("c_c_add", 1, 1, 2.0, True)]


def _zdigest(z):
    return hashlib.sha256(np.asfortranarray(z, np.complex128).tobytes(order="F")).hexdigest()


def test_complex_gemm_bit_exact_with_tmv_block_path(orc, ref_vectors):
    """On an aligned shape TMV takes its block path, which splits a complex
    product into four real ones (src/TMV_MultMM_Block.cpp:370-392): the
    restatement (four real oracle GEMMs, same signs and order) reproduces
    TMV's bytes for every conjugation and alpha kind."""
    A = rng.uniform_matrix(SEED + 111, 128, 160) + 1j * rng.uniform_matrix(SEED + 112, 128, 160)
    B = rng.uniform_matrix(SEED + 113, 160, 96) + 1j * rng.uniform_matrix(SEED + 114, 160, 96)
    C0 = rng.uniform_matrix(SEED + 115, 128, 96) + 1j * rng.uniform_matrix(SEED + 116, 128, 96)
    for tag, ca, cb, alpha, add in ZGEMM_VARIANTS:
        Z = orc.gemm_complex(A, B, C=C0, alpha=alpha, add=add, conj_a=ca, conj_b=cb)
        assert _zdigest(Z) == str(ref_vectors[f"zgemm_128x96x160_{tag}_sha256"]), tag


def test_complex_gemm_ragged_close_to_tmv(orc, ref_vectors):
    """Ragged shapes go through TMV's small complex kernels instead (complex
    multiply-adds): the four-product form agrees to rounding."""
    A = rng.uniform_matrix(SEED + 121, 37, 41) + 1j * rng.uniform_matrix(SEED + 122, 37, 41)
    B = rng.uniform_matrix(SEED + 123, 41, 29) + 1j * rng.uniform_matrix(SEED + 124, 41, 29)
    Z = orc.gemm_complex(A, B, alpha=0.3 + 0.7j, conj_a=True)
    ref = ref_vectors["zgemm_37x29x41_conjA"]
    mag = np.abs(0.3 + 0.7j) * ((np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag)))
    assert np.all(np.abs(Z - ref) <= 4 * 41 * EPS64 * mag)


@pytest.mark.parametrize("ca,cb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_complex_gemm_integer_exact(orc, ca, cb):
    """Integer parts: every product and sum is exact, so the result equals
    numpy's complex product exactly; a dropped conjugation or a swapped
    sign fails."""
    g = np.random.default_rng(7 + 2 * ca + cb)
    A = g.integers(-9, 10, (23, 31)) + 1j * g.integers(-9, 10, (23, 31))
    B = g.integers(-9, 10, (31, 17)) + 1j * g.integers(-9, 10, (31, 17))
    C0 = g.integers(-9, 10, (23, 17)) + 1j * g.integers(-9, 10, (23, 17))
    Aop = A.conj() if ca else A
    Bop = B.conj() if cb else B
    for alpha, add in ((1.0, False), (2.0 - 3.0j, True)):
        Z = orc.gemm_complex(A, B, C=C0, alpha=alpha, add=add, conj_a=ca, conj_b=cb)
        assert np.array_equal(Z, alpha * (Aop @ Bop) + (C0 if add else 0))
    Zw = orc.gemm_complex(A, B, conj_a=not ca, conj_b=cb)
    assert not np.array_equal(Zw, Aop @ Bop)


def _zmat(seed, n):
    return rng.uniform_matrix(seed, n, n) + 1j * rng.uniform_matrix(seed + 10, n, n)


def _zresidual(A, LU, P):
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = np.array(A, np.complex128)
    for i, p in enumerate(P):
        if p != i:
            PA[[i, p]] = PA[[p, i]]
    return float(np.abs(PA - L @ U).max() / np.abs(A).max())


def _gepp_modulus(A):
    """Textbook right-looking GEPP in numpy with TMV's complex pivot rule
    (largest modulus, first row on ties).  LAPACK's zgetrf ranks by
    |re| + |im| instead, so it is no pin for complex pivots."""
    A = np.array(A, np.complex128)
    n = A.shape[0]
    P = np.zeros(n, np.int64)
    for j in range(n):
        p = j + int(np.argmax(np.abs(A[j:, j])))
        P[j] = p
        A[[j, p]] = A[[p, j]]
        A[j + 1:, j] *= 1.0 / A[j, j]
        A[j + 1:, j + 1:] -= np.outer(A[j + 1:, j], A[j, j + 1:])
    return A, P


@pytest.mark.parametrize("n", [5, 37, 100, 257])
def test_complex_lu_pivots_match_tmv(orc, ref_vectors, n):
    """Complex LU: the same pivot sequence as TMV's (fixtures) and as a
    textbook modulus-pivoting GEPP on tie-free inputs; P A = L U to 8 n eps."""
    A = rng.uniform_matrix(SEED + 230 + n, n, n) + 1j * rng.uniform_matrix(SEED + 240 + n, n, n)
    LU, P, gaps = orc.getrf_complex(A, gaps=True)
    assert np.array_equal(P, ref_vectors[f"zlu_{n}_P"])
    assert (gaps > 64 * n * EPS64).all()
    if n <= 100:
        _, Pb = _gepp_modulus(A)
        assert np.array_equal(P, Pb)
    assert _zresidual(A, LU, P) <= 8 * n * EPS64


def test_complex_lu_zero_column_and_solve(orc):
    """The underflow rule on a complex column, and the solves (both
    transposes) to 8 n eps backward error; a planted wrong pivot fails."""
    n = 60
    A = _zmat(SEED + 77, n)
    A[:, 7] = 0.0
    LU, P = orc.getrf_complex(A)
    assert P[7] == 7 and np.all(LU[7:, 7] == 0)
    B = _zmat(SEED + 88, n)
    A2 = _zmat(SEED + 99, n)
    LU2, P2 = orc.getrf_complex(A2)
    for trans in (False, True):
        X = orc.getrs_complex(LU2, P2, B, trans=trans)
        M = A2.T if trans else A2
        r = np.abs(B - M @ X).max() / (np.abs(M).max() * np.abs(X).max() * n)
        assert r <= 8 * EPS64
    Pw = P2.copy()
    Pw[3] = (Pw[3] + 1) % n if Pw[3] + 1 < n else 3
    assert _zresidual(A2, LU2, Pw) > 1e-3


# ------------------------------------------------------------------ QR ----

def _q_from(QR, beta):
    """Q = H_0 H_1 ... H_{n-1} applied to the identity (m x m)."""
    m, n = QR.shape
    Q = np.eye(m)
    for j in reversed(range(n)):
        v = np.concatenate(([1.0], QR[j + 1:, j]))
        Q[j:] -= beta[j] * np.outer(v, v @ Q[j:])
    return Q


@pytest.mark.parametrize("m,n", [(37, 37), (100, 64)])
def test_qr_matches_tmv_and_lapack(orc, ref_vectors, m, n):
    """The restatement of TMV's NonBlockQRDecompose + HouseholderReflect:
    beta and R equal TMV's (fixtures) and LAPACK dgeqrf's (the same
    reflector convention) to rounding; A = Q R to 8 m eps; QR division
    equals TMV's x = b/A and numpy's least squares."""
    A = rng.uniform_matrix(SEED + 260 + m + n, m, n)
    QR, beta = orc.qr_decompose(A)
    tol = 64 * m * EPS64 * np.abs(A).max()
    assert np.abs(beta - ref_vectors[f"qr_{m}x{n}_beta"]).max() <= 64 * m * EPS64
    assert np.abs(np.triu(QR[:n]) - ref_vectors[f"qr_{m}x{n}_R"]).max() <= tol
    lq, ltau = sla.lapack.dgeqrf(A)[:2]
    assert np.abs(np.triu(QR[:n]) - np.triu(lq[:n])).max() <= tol
    assert np.abs(beta - ltau).max() <= 64 * m * EPS64
    Q = _q_from(QR, beta)
    R = np.zeros((m, n))
    R[:n] = np.triu(QR[:n])
    assert np.abs(A - Q @ R).max() <= 8 * m * EPS64 * np.abs(A).max()
    B = rng.uniform_matrix(SEED + 270 + m, m, 2)
    X = orc.qr_solve(QR, beta, B)
    np.testing.assert_allclose(X, ref_vectors[f"qr_{m}x{n}_x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(X, np.linalg.lstsq(A, B, rcond=None)[0], rtol=0, atol=1e-12)
    # a reflector with the wrong sign convention breaks the reconstruction
    Rw = R.copy()
    Rw[0] = -Rw[0]
    assert np.abs(A - Q @ Rw).max() > 1e-3


def test_qr_zero_column_and_singular(orc):
    """A column already zero below (and on) the diagonal gets beta = 0 and
    R(j, j) = 0 (HouseholderReflect's no-reflection cases); QR division then
    reports a singular R."""
    A = rng.uniform_matrix(SEED + 281, 40, 12)
    A[:, 5] = 0.0
    QR, beta = orc.qr_decompose(A)
    ref = orc.ref_geqrf(A) if orc.ref_lib() is not None else None
    assert QR[5, 5] == 0.0
    if ref is not None:
        assert abs(ref[0][5, 5]) <= 1e-12 and abs(beta[5] - ref[1][5]) <= 1e-12
    with pytest.raises(orc.SingularError):
        orc.qr_solve(QR, beta, np.ones(40))


def test_inverse_example_golden(orc):
    """examples/Division.cpp:75-133: A2.inverse(), 5./A2, A2.inverse() * b
    and b * A2.inverse() to the 6 printed significant digits."""
    A2, b = F.inverse_example()
    K = F.KNOWN["inverse_example"]
    LU, P = orc.getrf(A2)
    X = orc.inverse_from_lu(LU, P)
    np.testing.assert_allclose(X, K["A2inv"], rtol=5e-6)
    np.testing.assert_allclose(5.0 * X, K["five_over_A2"], rtol=5e-6)
    np.testing.assert_allclose(X @ b, K["inv_times_b"], rtol=5e-6)
    np.testing.assert_allclose(b @ X, K["b_times_inv"], rtol=5e-6)
    # a transposed or unpermuted inverse fails the fixture
    assert not np.allclose(X.T, K["A2inv"], rtol=5e-6)
    assert not np.allclose(orc.getrs(LU, np.arange(4), np.eye(4)), K["A2inv"], rtol=5e-6)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 37, 64, 100, 130])
def test_inverse_matches_tmv_ref_vectors(orc, ref_vectors, n):
    """The solve of A X = I equals TMV's m.inverse() (LU_Inverse's
    U^-1 L^-1 P^T) to A's conditioning, and A X = I to a few eps."""
    M = rng.uniform_matrix(SEED + 200 + n, n, n)
    LU, P = orc.getrf(M)
    X = orc.inverse_from_lu(LU, P)
    ok, r = inverse_close(X, ref_vectors[f"inv_{n}"], M, EPS64)
    assert ok, r
    assert inverse_residual(M, X) <= 8 * EPS64
    assert inverse_residual(M, ref_vectors[f"inv_{n}"]) <= 8 * EPS64


@pytest.mark.parametrize("n", [1, 2, 3, 5, 37, 64, 100, 130])
def test_logdet_matches_tmv_ref_vectors(orc, ref_vectors, n):
    """logDet(&sign) from the factors equals TMV's; independently,
    LAPACK's slogdet (numpy) agrees."""
    M = rng.uniform_matrix(SEED + 200 + n, n, n)
    LU, P = orc.getrf(M)
    ld, sg = orc.logdet_from_lu(LU, P)
    ref_ld, ref_sg = ref_vectors[f"logdet_{n}"]
    assert sg == ref_sg
    assert abs(ld - ref_ld) <= 4 * n * EPS64 * max(1.0, abs(ref_ld))
    np_sg, np_ld = np.linalg.slogdet(M)
    assert sg == np_sg and abs(ld - np_ld) <= 1e-10 * max(1.0, abs(np_ld))


def test_logdet_fixtures_and_singular(orc, ref_vectors):
    """sign * exp(logdet) reproduces TMV's exact determinants
    (tests/TMV_TestMatrixDet.cpp); the cfg1 matrix A + 512 I (det far beyond
    double range) matches TMV's logDet; a zero pivot gives (-inf, 0) as
    DiagMatrix::logDet does (src/TMV_DiagMatrix.cpp:67-69)."""
    mats = F.det_matrices()
    for name, (true, k) in F.KNOWN["dets"].items():
        m = mats[name]
        LU, P = orc.getrf(m)
        ld, sg = orc.logdet_from_lu(LU, P)
        d = 0.0 if sg == 0 else sg * np.exp(ld)
        norm_m = mats["m11"] if name == "m12" else m
        assert abs(d - true) <= k * float(np.sum(norm_m ** 2)) * EPS, name
    n = 512
    As = rng.uniform_matrix(SEED + 0, n, n) + n * np.eye(n)
    LU, P = orc.getrf(As)
    ld, sg = orc.logdet_from_lu(LU, P)
    assert sg == ref_vectors["cfg1_logdet"][1]
    assert abs(ld - ref_vectors["cfg1_logdet"][0]) <= 4 * n * EPS64 * ld
    A = F.decomp_matrix(3, N=20)
    LU, P = orc.getrf(A)
    ld, sg = orc.logdet_from_lu(LU, P)
    assert sg == 0.0 and ld == -np.inf
    if orc.ref_lib() is not None:
        assert orc.ref_logdet(A) == (-np.inf, 0.0)


def test_division_example_golden(orc):
    """examples/Division.cpp:32-54 to the 6 printed significant digits."""
    A, b = F.division_example()
    K = F.KNOWN["division_example"]
    x = orc.gesv(A, b)
    np.testing.assert_allclose(x, K["x"], rtol=5e-6)
    A2 = A.copy()
    A2[0, 0] = 20.0
    np.testing.assert_allclose(orc.gesv(A2, b), K["x_after_A00_20"], rtol=5e-6)
    np.testing.assert_allclose(orc.gesv(A2, b[::-1].copy()), K["x2_reversed_b"], rtol=5e-6)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 65, 100, 130, 257])
def test_lu_pivots_match_lapack_and_reconstruct(orc, n):
    """Tie-free GEPP pivots are fixed by the mathematics: LAPACK getrf
    (scipy) must pick the same rows; P A = L U to 8 n eps."""
    A = rng.uniform_matrix(SEED + 400 + n, n, n)
    LU, P, gaps = orc.getrf(A, gaps=True)
    lu, piv = sla.lu_factor(A)
    tie_free = gaps > 64 * n * EPS64
    assert tie_free.all()
    assert np.array_equal(P, piv)
    assert lu_residual(A, LU, P) <= 8 * n * EPS64


def test_lu_matches_tmv_ref_vectors(orc, ref_vectors):
    for n in (1, 2, 3, 5, 37, 64, 100, 130):
        A = rng.uniform_matrix(SEED + 200 + n, n, n)
        LU, P = orc.getrf(A)
        assert np.array_equal(P, ref_vectors[f"lu_{n}_P"]), n
        np.testing.assert_allclose(LU, ref_vectors[f"lu_{n}_LU"], rtol=0,
                                   atol=64 * n * EPS64 * np.abs(ref_vectors[f"lu_{n}_LU"]).max())
        b = rng.uniform(SEED + 300 + n, n, -1.0, 1.0)
        x = orc.getrs(LU, P, b)
        xr = ref_vectors[f"solve_{n}_x"]
        assert np.abs(x - xr).max() <= 1e3 * n * EPS64 * max(1.0, np.abs(xr).max()), n


@pytest.mark.parametrize("mattype", [0, 3, 4, 5, 6])
def test_decomp_fixtures(orc, ref_vectors, mattype):
    """tests/TMV_TestMatrixDecomp.cpp matrices: regular, singular, bad
    defects, near underflow (TMV's zero-column rule), near overflow."""
    A = F.decomp_matrix(mattype)
    LU, P, gaps = orc.getrf(A, gaps=True)
    Pr = ref_vectors[f"decomp_{mattype}_P"]
    # Pivots agree up to the first column whose choice is not clear-cut
    # (singular inputs end in rounding noise, where any order is valid).
    amb = np.nonzero(gaps <= 64 * 200 * EPS64)[0]
    k0 = int(amb[0]) if len(amb) else 200
    assert np.array_equal(P[:k0], Pr[:k0])
    if mattype == 0:
        assert k0 == 200
        LUr = ref_vectors["decomp_0_LU"]
        assert np.abs(LU - LUr).max() <= 64 * 200 * EPS64 * np.abs(LUr).max()
    if mattype in (0, 3, 4):
        assert lu_residual(A, LU, P) <= 8 * 200 * EPS64


def test_big_div_fixture(orc, ref_vectors):
    A = F.big_div_100()
    LU, P = orc.getrf(A)
    assert np.array_equal(P, ref_vectors["bigdiv_P"])
    assert lu_residual(A, LU, P) <= 8 * 100 * EPS64


def test_cfg1_lu_matches_tmv(orc, ref_vectors):
    n = 512
    A = rng.uniform_matrix(SEED + 0, n, n) + n * np.eye(n)
    LU, P = orc.getrf(A)
    assert np.array_equal(P, ref_vectors["cfg1_lu_P"])
    b = rng.uniform(SEED + 2, n, -1.0, 1.0)
    x = orc.getrs(LU, P, b)
    np.testing.assert_allclose(x, ref_vectors["cfg1_x"], rtol=0, atol=1e-14)


def test_lu_planted_wrong_pivot_is_caught(orc):
    """A mistake planted in the oracle's output (one interchange flipped)
    fails the reconstruction check."""
    n = 40
    A = rng.uniform_matrix(SEED + 500, n, n)
    LU, P = orc.getrf(A)
    P2 = P.copy()
    P2[3] = 3 if P[3] != 3 else 4
    assert lu_residual(A, LU, P2) > 8 * n * EPS64


def test_solve_transposed(orc):
    n = 70
    A = rng.uniform_matrix(SEED + 501, n, n)
    b = rng.uniform(SEED + 502, n, -1.0, 1.0)
    LU, P = orc.getrf(A)
    x = orc.getrs(LU, P, b, trans=True)
    assert backward_error(A.T, x, b) <= 8 * n * EPS64


def test_singular_solve_raises(orc):
    """A zero U diagonal met by a non-zero entry raises (TMV throws
    SingularUpperTriMatrix, src/TMV_TriDiv_V.cpp:171-217)."""
    A = F.decomp_matrix(3, N=20)
    LU, P = orc.getrf(A)
    with pytest.raises(orc.SingularError):
        orc.getrs(LU, P, np.ones(20))


def test_float32_lu(orc):
    n = 96
    A = rng.uniform_matrix(SEED + 503, n, n)
    LU, P = orc.getrf(A, dtype=np.float32)
    assert lu_residual(A.astype(np.float32), LU, P) <= 8 * n * EPS32


def test_rng_reference_values():
    """The documented generator: fixed words for seed 0 (SplitMix64)."""
    w = rng.words(0, 0, 3)
    k = rng.key(0)
    assert int(k) == 0xE220A8397B1DCDAF
    assert w.dtype == np.uint64 and len(set(w.tolist())) == 3
    u = rng.uniform(SEED, 1000, -1.0, 1.0)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(u.mean()) < 0.1


# ---------------------------------------------------- live real-TMV ---

def test_live_reference_agrees(orc):
    """Where oracle/_ref was built, compare the restatement with TMV live."""
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built here")
    A = rng.uniform_matrix(SEED + 600, 96, 96)
    B = rng.uniform_matrix(SEED + 601, 96, 96)
    assert np.array_equal(orc.gemm(A, B), orc.gemm(A, B, lib=ref))
    LU, P = orc.getrf(A)
    LU2, P2 = orc.getrf(A, lib=ref)
    assert np.array_equal(P, P2)


# ------------------------------------------------- exact ties, triangles ----

@pytest.mark.parametrize("n", [64, 256, 1024])
def test_hadamard_ties_match_tmv(orc, ref_vectors, n):
    """Row-shuffled Sylvester-Hadamard matrices: every elimination step is an
    exact tie (gap 0) and the arithmetic is exact, so TMV's first-index rule
    (src/TMV_Vector.cpp:473-504) alone fixes the pivots.  The restatement
    gives TMV's pivots and bit-identical factors; a last-index rule does
    not."""
    import make_ref_vectors as M
    H = M.hadamard_rows(n)
    LU, P, gaps = orc.getrf(H, gaps=True)
    assert np.array_equal(P, ref_vectors[f"hadamard_{n}_P"])
    assert hashlib.sha256(np.asfortranarray(LU).tobytes(order="F")).hexdigest() == \
        str(ref_vectors[f"hadamard_{n}_LU_sha256"])
    assert (gaps[:-1] == 0).all()                      # ties at every step
    # a planted last-index rule chooses differently on the first tie
    col = np.abs(H[:, 0])
    assert int(np.nonzero(col == col.max())[0][-1]) != int(P[0])


def test_cfg2_fixture_is_tie_free(ref_vectors):
    """The stored n = 8192 pivot sequence is tie-free well beyond 64 n eps,
    so GPU pivots must equal it at every step."""
    n = 8192
    assert ref_vectors["cfg2_lu_P"].shape == (n,)
    assert float(ref_vectors["cfg2_lu_min_gap"]) > 1e3 * 64 * n * EPS64


def _tri(n, upper, seed, zero_diag=None):
    A = rng.uniform_matrix(seed, n, n) / max(n, 1)
    d = 1.0 + rng.uniform(seed + 1, n, 0.0, 1.0)
    A[np.diag_indices(n)] = d * np.where(rng.uniform(seed + 2, n) < 0.5, -1.0, 1.0)
    A = np.triu(A) if upper else np.tril(A)
    if zero_diag is not None:
        A[zero_diag, zero_diag] = 0.0
    return A


@pytest.mark.parametrize("upper", [0, 1])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("unit", [0, 1])
def test_trsm_vs_lapack(orc, upper, trans, unit):
    """The restated recursive TRSM (src/TMV_TriDiv_M.cpp:190-266) against
    LAPACK's trsm (scipy) for every triangle/transpose/unit combination; a
    solve with the wrong triangle fails the same check."""
    n, nrhs = 150, 7
    A = _tri(n, upper, SEED + 700 + 4 * upper + 2 * trans + unit)
    B = rng.uniform_matrix(SEED + 710, n, nrhs)
    X = orc.trsm(A, B, upper=upper, trans=trans, unit=unit)
    Xl = sla.solve_triangular(A, B, lower=not upper, trans=trans, unit_diagonal=bool(unit))
    assert np.abs(X - Xl).max() <= 64 * n * EPS64 * np.abs(Xl).max()
    Xw = orc.trsm(A.T, B, upper=not upper, trans=trans, unit=unit)     # wrong triangle
    assert np.abs(Xw - Xl).max() > 1e-3 * np.abs(Xl).max()


def test_trsm_zero_diagonal(orc):
    """A zero diagonal met by a non-zero entry raises (TMV's
    SingularUpperTriMatrix); met only by zero entries it does not
    (src/TMV_TriDiv_V.cpp:171-217 skips zero entries)."""
    n = 40
    A = _tri(n, 1, SEED + 720, zero_diag=10)
    with pytest.raises(orc.SingularError):
        orc.trsm(A, rng.uniform(SEED + 721, n, -1, 1), upper=1)
    b = np.zeros(n)
    b[:5] = 1.0                                    # rows 0..4 only: row 10 stays 0
    orc.trsm(A, b, upper=1)


# ------------------------------------------------ non-square LU views ---

RECT = [(5, 3), (3, 5), (100, 37), (37, 100), (300, 130), (130, 300), (1, 7), (7, 1)]


@pytest.mark.parametrize("m,n", RECT)
def test_getrf_rect_matches_tmv(orc, ref_vectors, m, n):
    """The restatement on non-square matrices (R = min(m, n) pivots,
    src/TMV_LUDecompose.cpp:217-245): TMV's pivots exactly and TMV's factors
    to rounding.  P A = L U holds on the first min(m, n) columns; for a wide
    matrix TMV's native path leaves the columns past m off L^-1 P A (its
    dgetrf seam, src/TMV_LUDecompose.cpp:418-442, does not), and the
    restatement follows TMV there too."""
    from checks import EPS64, lu_residual_mn
    A = rng.uniform_matrix(SEED + 700 + m + 3 * n, m, n)
    LU, P = orc.getrf(A)
    assert np.array_equal(P, ref_vectors[f"lurect_{m}x{n}_P"])
    LUt = ref_vectors[f"lurect_{m}x{n}_LU"]
    bound = 8 * max(m, n) * EPS64
    r = min(m, n)
    assert lu_residual_mn(A[:, :r], LU[:, :r], P) <= bound
    assert lu_residual_mn(A[:, :r], LUt[:, :r], P) <= bound
    assert np.abs(LU - LUt).max() <= bound * max(1.0, np.abs(LUt).max())
    if m >= n:
        assert lu_residual_mn(A, LU, P) <= bound


def test_lu_residual_mn_catches_errors(orc):
    """The rectangular residual check fails on a planted mistake: a swapped
    interchange, or a wrong multiplier."""
    from checks import EPS64, lu_residual_mn
    A = rng.uniform_matrix(SEED + 705, 60, 25)
    LU, P = orc.getrf(A)
    assert lu_residual_mn(A, LU, P) <= 8 * 60 * EPS64
    Pb = P.copy()
    Pb[3] = (Pb[3] + 7) % 60 if Pb[3] + 7 >= 3 else Pb[3] + 7
    assert lu_residual_mn(A, LU, Pb) > 1e-3
    LUb = LU.copy()
    LUb[40, 10] += 1e-6
    assert lu_residual_mn(A, LUb, P) > 8 * 60 * EPS64
