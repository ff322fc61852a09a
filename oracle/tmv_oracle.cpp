// tmv_oracle.cpp -- CPU restatement of TMV 0.76's dense GEMM + LU path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
// check and bench.py's CPU-baseline leg may load it; the product library
// (tmv_amd/libtmvx.so) never links, loads or falls back to it.
//
// It restates, in fresh code, the algorithms of the reference files cited at
// each function (paths relative to the TMV source tree).  Every function is
// instantiated for float, double and long double; the long double build is
// the cross-check used for forward errors.
//
// Pinning (what checks this file; see tests/test_oracle_*.py):
//   orc_gemm_*   integer-exact products (examples/Matrix.cpp:296-334),
//                long-double brute force, layout invariance, and the real TMV
//                library (oracle/_ref) where it builds.
//   orc_getrf_*  exact determinants (tests/TMV_TestMatrixDet.cpp:7-160,
//                tests/TMV_TestMatrixDiv.cpp:11-18), reconstruction
//                ||A - P^T L U||, pivot sequence vs LAPACK getrf (scipy),
//                singular / underflow fixtures, and oracle/_ref bit-for-bit.
//   orc_getrs_*  golden solve examples/Division.cpp:22-73, backward error,
//                and oracle/_ref.
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstring>
#include <limits>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef ptrdiff_t idx;

// Real type of an element type and TMV_ABS (std::abs: the modulus for
// complex elements, include/tmv/TMV_Base.h).
template <class T> struct Real { typedef T type; };
template <class R> struct Real<std::complex<R>> { typedef R type; };
template <class T> typename Real<T>::type oabs(T x) { return std::abs(x); }

// A strided matrix view: element (i,j) at p[i*si + j*sj].
template <class T> struct View {
    T* p; idx rows, cols, si, sj;
    T& operator()(idx i, idx j) const { return p[i * si + j * sj]; }
    View sub(idx i0, idx i1, idx j0, idx j1) const {
        View v = {p + i0 * si + j0 * sj, i1 - i0, j1 - j0, si, sj};
        return v;
    }
};

template <class T> View<T> make_view(T* p, idx rows, idx cols, idx ld, int rm)
{
    View<T> v = {p, rows, cols, rm ? ld : 1, rm ? 1 : ld};
    return v;
}

// Block sizes of BlockHelper<T> (src/TMV_MultMM_Block.cpp:119-164):
// 16x16 C tiles, K blocks of 32 (double) or 64 (float).
template <class T> struct KBlock { enum { value = 32 }; };
template <> struct KBlock<float> { enum { value = 64 }; };

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------

// One 16x16 (or edge) C tile over all of K, following SingleBlockMultMM
// (src/TMV_MultMM_Block.cpp:35-94) with the micro-kernel's summation order
// (src/TMV_MultMM_Kernel.h:2318-2512): inside each full K block, two partial
// sums -- even k and odd k, each a running sum of products in k order -- are
// added together and then into the tile scratch Cp.  The K tail
// (call_multmm_16_16_K, Block.cpp:72-82) is a running sum in k order.
// Finally C = alpha*Cp or C += alpha*Cp (Block.cpp:84-93).
template <class T>
void block_tile(bool add, T alpha, const View<T>& A, const View<T>& B,
                const View<T>& C, idx i1, idx i2, idx j1, idx j2)
{
    const idx K = A.cols;
    const idx KB = KBlock<T>::value;
    const idx Kfull = (K / KB) * KB;
    T Cp[16][16];
    for (idx i = 0; i < 16; ++i)
        for (idx j = 0; j < 16; ++j) Cp[i][j] = T(0);
    for (idx i = i1; i < i2; ++i) {
        for (idx j = j1; j < j2; ++j) {
            T acc = T(0);
            // The kernel walks rows in groups of four: rows 0,1 of a group
            // run k ascending, rows 2,3 run k descending (Kernel.h:2329-2500).
            const bool down = ((i - i1) & 3) >= 2;
            for (idx k0 = 0; k0 < Kfull; k0 += KB) {
                T e, o;
                if (!down) {
                    e = A(i, k0) * B(k0, j);
                    o = A(i, k0 + 1) * B(k0 + 1, j);
                    for (idx k = k0 + 2; k < k0 + KB; k += 2) {
                        e = e + A(i, k) * B(k, j);
                        o = o + A(i, k + 1) * B(k + 1, j);
                    }
                } else {
                    const idx kl = k0 + KB - 2;
                    e = A(i, kl) * B(kl, j);
                    o = A(i, kl + 1) * B(kl + 1, j);
                    for (idx k = kl - 2; k >= k0; k -= 2) {
                        e = e + A(i, k) * B(k, j);
                        o = o + A(i, k + 1) * B(k + 1, j);
                    }
                }
                acc = acc + (e + o);
            }
            // K tail (sse2_multmm_16_16_K, Kernel.h:3164-3300): the same
            // even/odd pairing over the tail, ascending, zero-padded to even.
            if (Kfull < K) {
                T e = A(i, Kfull) * B(Kfull, j);
                T o = (Kfull + 1 < K) ? A(i, Kfull + 1) * B(Kfull + 1, j) : T(0);
                for (idx k = Kfull + 2; k < K; k += 2) {
                    e = e + A(i, k) * B(k, j);
                    if (k + 1 < K) o = o + A(i, k + 1) * B(k + 1, j);
                }
                acc = acc + (e + o);
            }
            Cp[i - i1][j - j1] = acc;
        }
    }
    for (idx j = j1; j < j2; ++j)
        for (idx i = i1; i < i2; ++i) {
            T v = alpha * Cp[i - i1][j - j1];
            C(i, j) = add ? C(i, j) + v : v;
        }
}

// Simple kernels for tiny/thin shapes (CCCMultMM / CRCMultMM,
// src/TMV_MultMM_CCC.cpp:33-138, src/TMV_MultMM_CRC.cpp:32-136): column axpys
// over ascending k into a temporary, then C (+)= alpha*temp.  RCCMultMM
// (src/TMV_MultMM_RCC.cpp:32-733) runs k descending (RCC.cpp:38).
template <class T>
void simple_mm(bool add, T alpha, const View<T>& A, const View<T>& B,
               const View<T>& C, bool a_rowmajor)
{
    const idx M = C.rows, N = C.cols, K = A.cols;
    std::vector<T> tmp((size_t)(M * N), T(0));
    for (idx j = 0; j < N; ++j)
        for (idx i = 0; i < M; ++i) {
            T acc = T(0);
            if (a_rowmajor) for (idx k = K - 1; k >= 0; --k) acc += A(i, k) * B(k, j);
            else for (idx k = 0; k < K; ++k) acc += A(i, k) * B(k, j);
            tmp[(size_t)(i + j * M)] = acc;
        }
    for (idx j = 0; j < N; ++j)
        for (idx i = 0; i < M; ++i) {
            T v = alpha * tmp[(size_t)(i + j * M)];
            C(i, j) = add ? C(i, j) + v : v;
        }
}

// MultMM<add> (src/TMV_MultMM.cpp:404-477) with NonBlasMultMM routing
// (src/TMV_MultMM.cpp:46-100).  C is column-major.  The OpenMP split
// (src/TMV_MultMM_OpenMP.cpp:31-66) hands each thread slabs of whole 16-wide
// tiles, so results do not depend on the thread count; here the tiles are
// shared out by an OpenMP loop.
template <class T>
void multmm(bool add, T alpha, const View<T>& A, const View<T>& B,
            const View<T>& C, bool a_rowmajor)
{
    const idx M = C.rows, N = C.cols, K = A.cols;
    if (M == 0 || N == 0) return;
    if (K == 0 || alpha == T(0)) {
        if (!add)
            for (idx j = 0; j < N; ++j)
                for (idx i = 0; i < M; ++i) C(i, j) = T(0);
        return;
    }
    const idx Mb = M >> 6, Nb = N >> 6, Kb = K >> 6;
    const idx Mc = M < 16 ? 1 : (M >> 4), Nc = N < 16 ? 1 : (N >> 4),
              Kc = K < 16 ? 1 : (K >> 4);
    const bool twobig = (Mb && Nb) || (Mb && Kb) || (Nb && Kb);
    if ((M < 16 && N < 16 && K < 16) || (M <= 3 || N <= 3 || K <= 3) ||
        ((M < 16 || N < 16 || K < 16) && (!twobig || (Mc * Nc * Kc < 4)))) {
        simple_mm(add, alpha, A, B, C, a_rowmajor);
        return;
    }
    const idx ntm = (M + 15) / 16, ntn = (N + 15) / 16;
#pragma omp parallel for schedule(static) collapse(2)
    for (idx tj = 0; tj < ntn; ++tj)
        for (idx ti = 0; ti < ntm; ++ti)
            block_tile(add, alpha, A, B, C, ti * 16, std::min(M, ti * 16 + 16),
                       tj * 16, std::min(N, tj * 16 + 16));
}

// ---------------------------------------------------------------------------
// Pivot primitives
// ---------------------------------------------------------------------------

// FindMaxAbsElement (src/TMV_Vector.cpp:473-504): first index wins ties.
template <class T> typename Real<T>::type max_abs(const T* v, idx n, idx step, idx* imax)
{
    typedef typename Real<T>::type RT;
    RT best = RT(-1);
    *imax = 0;
    for (idx i = 0; i < n; ++i) {
        RT a = oabs(v[i * step]);
        if (a > best) { best = a; *imax = i; }
    }
    return n > 0 ? best : RT(0);
}

// TMV_Underflow (include/tmv/TMV_Base.h:692-697): |x| < numeric min.
template <class RT> bool underflow(RT x)
{ return std::fabs(x) < std::numeric_limits<RT>::min(); }

// swapRows over a column range (src/TMV_Matrix.cpp:865-930 apply the
// interchange list in order; block width does not change the result).
template <class T> void swap_rows(const View<T>& A, idx r1, idx r2)
{
    if (r1 == r2) return;
    for (idx j = 0; j < A.cols; ++j) std::swap(A(r1, j), A(r2, j));
}

template <class T> void permute_rows(const View<T>& A, const idx* P, idx i1, idx i2)
{
    for (idx i = i1; i < i2; ++i) swap_rows(A, i, P[i]);
}

// Optional record of pivot gaps: for each pivot search, the relative gap
// (|best| - |runner-up|) / |best| among the candidates.  Not part of TMV;
// lets tests assert pivot equality only on tie-free inputs.
struct GapLog { double* out; idx n; };

template <class T>
void log_gap(GapLog* g, idx col, const T* v, idx n, idx step, typename Real<T>::type best)
{
    typedef typename Real<T>::type RT;
    if (!g || !g->out || col < 0 || col >= g->n) return;
    if (!(best > RT(0))) { g->out[col] = 0.0; return; }
    RT second = RT(0);
    bool seen = false;
    for (idx i = 0; i < n; ++i) {
        RT a = oabs(v[i * step]);
        if (a == best && !seen) { seen = true; continue; }
        if (a > second) second = a;
    }
    g->out[col] = (double)((best - second) / best);
}

// ---------------------------------------------------------------------------
// Triangular solves
// ---------------------------------------------------------------------------

// Vector solve, lower triangle, column sweep (src/TMV_TriDiv_V.cpp:321-383)
// after skipping leading zeros of b (:440-461).  Returns -2 on a zero
// diagonal met by a non-zero b_j (SingularLowerTriMatrix).
template <class T> int trsv_lower(const View<T>& L, bool unit, T* b, idx bs)
{
    const idx N = L.rows;
    idx i1 = 0;
    while (i1 < N && b[i1 * bs] == T(0)) ++i1;
    for (idx j = i1; j < N; ++j) {
        if (b[j * bs] == T(0)) continue;
        if (!unit) {
            if (L(j, j) == T(0)) return -2;
            b[j * bs] = b[j * bs] / L(j, j);
        }
        const T bj = b[j * bs];
        for (idx i = j + 1; i < N; ++i) b[i * bs] -= bj * L(i, j);
    }
    return 0;
}

// Vector solve, upper triangle, column sweep (src/TMV_TriDiv_V.cpp:147-218)
// after skipping trailing zeros of b (:419-438).
template <class T> int trsv_upper(const View<T>& U, bool unit, T* b, idx bs)
{
    idx i2 = U.rows;
    while (i2 > 0 && b[(i2 - 1) * bs] == T(0)) --i2;
    for (idx j = i2 - 1; j >= 0; --j) {
        if (b[j * bs] == T(0)) continue;
        if (!unit) {
            if (U(j, j) == T(0)) return -2;
            b[j * bs] = b[j * bs] / U(j, j);
        }
        const T bj = b[j * bs];
        for (idx i = 0; i < j; ++i) b[i * bs] -= bj * U(i, j);
    }
    return 0;
}

// Matrix solve, NonBlasTriLDivEq (src/TMV_TriDiv_M.cpp:190-266): recursive
// halving with the split rounded down to a multiple of 64 and a MultMM
// update; base case N <= 32 solves column by column.
template <class T> int trsm_lower(const View<T>& L, bool unit, const View<T>& B)
{
    const idx N = L.rows;
    if (N <= 0 || B.cols <= 0) return 0;
    if (N <= 32) {
        for (idx j = 0; j < B.cols; ++j) {
            int rc = trsv_lower(L, unit, &B(0, j), B.si);
            if (rc) return rc;
        }
        return 0;
    }
    idx k = N / 2;
    if (k > 64) k = k / 64 * 64;
    int rc = trsm_lower(L.sub(0, k, 0, k), unit, B.sub(0, k, 0, B.cols));
    if (rc) return rc;
    multmm<T>(true, T(-1), L.sub(k, N, 0, k), B.sub(0, k, 0, B.cols),
              B.sub(k, N, 0, B.cols), false);
    return trsm_lower(L.sub(k, N, k, N), unit, B.sub(k, N, 0, B.cols));
}

template <class T> int trsm_upper(const View<T>& U, bool unit, const View<T>& B)
{
    const idx N = U.rows;
    if (N <= 0 || B.cols <= 0) return 0;
    if (N <= 32) {
        for (idx j = 0; j < B.cols; ++j) {
            int rc = trsv_upper(U, unit, &B(0, j), B.si);
            if (rc) return rc;
        }
        return 0;
    }
    idx k = N / 2;
    if (k > 64) k = k / 64 * 64;
    int rc = trsm_upper(U.sub(k, N, k, N), unit, B.sub(k, N, 0, B.cols));
    if (rc) return rc;
    multmm<T>(true, T(-1), U.sub(0, k, k, N), B.sub(k, N, 0, B.cols),
              B.sub(0, k, 0, B.cols), false);
    return trsm_upper(U.sub(0, k, 0, k), unit, B.sub(0, k, 0, B.cols));
}

// ---------------------------------------------------------------------------
// LU
// ---------------------------------------------------------------------------

// RecursiveLUDecompose (src/TMV_LUDecompose.cpp:217-407), column-major A
// (M x N), P gets min(M,N) 0-based interchanges.  Split at N1 = R/2 rounded
// down to a multiple of 64 when above 64; base cases R == 2 and R == 1 with
// the underflow rule (pivot |x| < min => zero the column, P[j] = j).
// Scaling by a pivot multiplies by its reciprocal (TMV_InverseOf,
// TMV_Base.h:719-720; VectorView /= x at TMV_VectorArith.h:105-107).
template <class T> void recursive_lu(const View<T>& A, idx* P, GapLog* g, idx col0)
{
    typedef typename Real<T>::type RT;
    const idx N = A.cols, M = A.rows, R = std::min(N, M);
    if (R > 2) {
        idx N1 = R / 2;
        if (N1 > 64) N1 = (N1 / 64) * 64;
        View<T> A0 = A.sub(0, M, 0, N1), A1 = A.sub(0, M, N1, N);
        recursive_lu(A0, P, g, col0);
        permute_rows(A1, P, 0, N1);
        trsm_lower(A.sub(0, N1, 0, N1), true, A.sub(0, N1, N1, N));
        multmm<T>(true, T(-1), A.sub(N1, M, 0, N1), A.sub(0, N1, N1, N),
                  A.sub(N1, M, N1, N), false);
        recursive_lu(A.sub(N1, M, N1, N), P + N1, g, col0 + N1);
        for (idx i = N1; i < R; ++i) P[i] += N1;
        permute_rows(A0, P, N1, R);
    } else if (R == 2) {
        T* a0 = &A(0, 0);
        T* a1 = &A(0, 1);
        idx ip0, ip1;
        RT piv = max_abs(a0, M, A.si, &ip0);
        log_gap(g, col0, a0, M, A.si, piv);
        if (underflow(piv)) {
            ip0 = 0; piv = RT(0);
            for (idx i = 0; i < M; ++i) a0[i * A.si] = T(0);
        }
        if (piv != RT(0)) {
            if (ip0 != 0) {
                std::swap(a0[ip0 * A.si], a0[0]);
                std::swap(a1[ip0 * A.si], a1[0]);
            }
            const T inv = T(1) / a0[0];
            const T a01 = a1[0];
            piv = RT(0);
            ip1 = 1;
            for (idx i = 1; i < M; ++i) {
                a0[i * A.si] *= inv;
                a1[i * A.si] -= a0[i * A.si] * a01;
                RT ab = oabs(a1[i * A.si]);
                if (ab > piv) { piv = ab; ip1 = i; }
            }
            log_gap(g, col0 + 1, a1 + A.si, M - 1, A.si, piv);
        } else {
            piv = max_abs(a1 + A.si, M - 1, A.si, &ip1);
            log_gap(g, col0 + 1, a1 + A.si, M - 1, A.si, piv);
            ++ip1;
        }
        if (underflow(piv)) {
            piv = RT(0); ip1 = 1;
            for (idx i = 1; i < M; ++i) a1[i * A.si] = T(0);
        }
        if (piv != RT(0) && M > 2) {
            if (ip1 != 1) {
                std::swap(a1[ip1 * A.si], a1[1 * A.si]);
                std::swap(a0[ip1 * A.si], a0[1 * A.si]);
            }
            const T inv = T(1) / a1[1 * A.si];
            for (idx i = 2; i < M; ++i) a1[i * A.si] *= inv;
        }
        if (N > 2) {
            // M == 2 < N: finish U(0:2, 2:N) (LUDecompose.cpp:345-353).
            if (ip0 == 1) swap_rows(A.sub(0, M, 2, N), 0, 1);
            const T a01 = a1[0];
            for (idx j = 2; j < N; ++j) A(1, j) -= a01 * A(0, j);
        }
        P[0] = ip0;
        P[1] = ip1;
    } else if (R == 1) {
        T* a0 = &A(0, 0);
        idx ip;
        RT piv = max_abs(a0, M, A.si, &ip);
        log_gap(g, col0, a0, M, A.si, piv);
        if (underflow(piv)) {
            piv = RT(0); ip = 0;
            for (idx i = 0; i < M; ++i) a0[i * A.si] = T(0);
        }
        if (piv != RT(0)) {
            if (ip != 0) std::swap(a0[ip * A.si], a0[0]);
            const T inv = T(1) / a0[0];
            for (idx i = 1; i < M; ++i) a0[i * A.si] *= inv;
        }
        P[0] = ip;
    }
}

// LU_LDivEq (src/TMV_LUDiv.cpp:185-226 with NonLapLULDivEq :46-55):
// permute the rows of B by P, then B /= L (unit), B /= U.
// Transposed (x^T A = b^T, i.e. A^T x = b), following LU_RDivEq
// (src/TMV_LUDiv.cpp:232-242, 369-410) on B^T: B^T %= U, B^T %= L, then
// reverse-permute the columns of B^T.  As column solves: U^T y = b,
// L^T z = y, then undo the interchanges in reverse order.
template <class T> int lu_solve(int trans, const View<T>& LU, const idx* P, const View<T>& B)
{
    const idx n = LU.rows;
    if (!trans) {
        permute_rows(B, P, 0, n);
        int rc = trsm_lower(LU, true, B);
        if (rc) return rc;
        return trsm_upper(LU, false, B);
    }
    View<T> LUt = {LU.p, LU.cols, LU.rows, LU.sj, LU.si};
    int rc = trsm_lower(LUt, false, B);   // U^T is lower, non-unit
    if (rc) return rc;
    rc = trsm_upper(LUt, true, B);        // L^T is upper, unit
    if (rc) return rc;
    for (idx i = n - 1; i >= 0; --i) swap_rows(B, i, P[i]);
    return 0;
}

template <class T>
int getrf_impl(idx m, idx n, T* A, idx lda, idx* P, double* gaps)
{
    if (m <= 0 || n <= 0) return 0;
    GapLog g = {gaps, std::min(m, n)};
    recursive_lu(make_view(A, m, n, lda, 0), P, gaps ? &g : 0, 0);
    return 0;
}

} // namespace

extern "C" {

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

#define ORC_GEMM(SUF, T)                                                      \
    int orc_gemm_##SUF(int a_rm, int b_rm, idx m, idx n, idx k, T alpha,      \
                       const T* A, idx lda, const T* B, idx ldb, int add,     \
                       T* C, idx ldc)                                         \
    {                                                                         \
        View<T> Av = make_view(const_cast<T*>(A), m, k, lda, a_rm);          \
        View<T> Bv = make_view(const_cast<T*>(B), k, n, ldb, b_rm);          \
        View<T> Cv = make_view(C, m, n, ldc, 0);                              \
        multmm<T>(add != 0, alpha, Av, Bv, Cv, a_rm != 0);                    \
        return 0;                                                             \
    }
ORC_GEMM(f64, double)
ORC_GEMM(f32, float)
ORC_GEMM(f80, long double)

#define ORC_LU(SUF, T)                                                        \
    int orc_getrf_##SUF(idx m, idx n, T* A, idx lda, idx* P, double* gaps)    \
    { return getrf_impl<T>(m, n, A, lda, P, gaps); }                          \
    int orc_getrs_##SUF(int trans, idx n, idx nrhs, const T* LU, idx lda,     \
                        const idx* P, T* B, idx ldb)                          \
    {                                                                         \
        if (n <= 0 || nrhs <= 0) return 0;                                    \
        return lu_solve<T>(trans, make_view(const_cast<T*>(LU), n, n, lda, 0), \
                           P, make_view(B, n, nrhs, ldb, 0));                 \
    }                                                                         \
    int orc_trsm_##SUF(int upper, int unit, idx n, idx nrhs, const T* A,      \
                       idx lda, T* B, idx ldb)                                \
    {                                                                         \
        View<T> Av = make_view(const_cast<T*>(A), n, n, lda, 0);             \
        View<T> Bv = make_view(B, n, nrhs, ldb, 0);                           \
        return upper ? trsm_upper<T>(Av, unit != 0, Bv)                       \
                     : trsm_lower<T>(Av, unit != 0, Bv);                      \
    }
ORC_LU(f64, double)
ORC_LU(f32, float)
ORC_LU(f80, long double)

// Complex LU and solve (std::complex<double>, interleaved storage): the same
// recursion with TMV_ABS = the modulus for the pivot search.  Its GEMM steps
// multiply complex elements directly, where TMV's block path splits them into
// four real products (src/TMV_MultMM_Block.cpp:370-392): pivots agree, the
// factors to rounding.
int orc_getrf_z(idx m, idx n, double* A, idx lda, idx* P, double* gaps)
{ return getrf_impl<std::complex<double>>(m, n, reinterpret_cast<std::complex<double>*>(A), lda, P, gaps); }
int orc_getrs_z(int trans, idx n, idx nrhs, const double* LU, idx lda, const idx* P, double* B,
                idx ldb)
{
    typedef std::complex<double> Z;
    if (n <= 0 || nrhs <= 0) return 0;
    return lu_solve<Z>(trans, make_view(reinterpret_cast<Z*>(const_cast<double*>(LU)), n, n, lda, 0),
                       P, make_view(reinterpret_cast<Z*>(B), n, nrhs, ldb, 0));
}

// Determinant from the factors: sign(P) * prod(diag U), the value
// LUDiv::det returns (src/TMV_LUD.cpp:132-156; Permutation::det at
// include/tmv/TMV_Permutation.h:273-278).  Computed in long double.
double orc_det_from_lu_f64(idx n, const double* LU, idx lda, const idx* P)
{
    long double d = 1.0L;
    for (idx i = 0; i < n; ++i) {
        d *= (long double)LU[i + i * lda];
        if (P[i] != i) d = -d;
    }
    return (double)d;
}

// log|det| and sign from the factors, the pair LUDiv::logDet returns
// (src/TMV_LUD.cpp:145-156): DiagMatrix::logDet (src/TMV_DiagMatrix.cpp:
// 59-90) sums log|U_ii| and multiplies the signs, a zero diagonal setting
// log|det| = log(0) = -inf and sign 0; det(P) flips the sign once per
// interchange.  Summed in long double.
double orc_logdet_from_lu_f64(idx n, const double* LU, idx lda, const idx* P, double* sign)
{
    long double sum = 0.0L;
    double s = 1.0;
    bool zero = false;
    for (idx i = 0; i < n; ++i) {
        const double d = LU[i + i * lda];
        if (d == 0.0) zero = true;
        else {
            sum += logl(fabsl((long double)d));
            if (d < 0.0) s = -s;
        }
        if (P[i] != i) s = -s;
    }
    *sign = zero ? 0.0 : s;
    return zero ? -INFINITY : (double)sum;
}

} // extern "C"
