// TMV-style tests of the drop-in C++ layer (include/tmvx_tmv.h) on the GPU.
// They follow TMV's own tests and examples (tests/TMV_TestMatrixDiv.cpp,
// tests/TMV_TestMatrixDet.cpp, examples/Division.cpp, examples/Matrix.cpp):
// the same matrices, the same division syntax, the same expected values.
// Exit status 0 when every check passes.
#include "tmvx_tmv.h"

#include <cmath>
#include <cstdio>

static int nfail = 0;
#define Assert(x, s)                                                           \
    do {                                                                       \
        if (!(x)) { std::printf("FAILED: %s (%s:%d)\n", s, __FILE__, __LINE__); ++nfail; } \
        else std::printf("ok: %s\n", s);                                       \
    } while (0)

static const double EPS = 10 * std::numeric_limits<double>::epsilon();

template <class T, int S>
static double NormF(const tmv::Matrix<T, S>& m)
{
    double s = 0;
    for (ptrdiff_t j = 0; j < m.rowsize(); ++j)
        for (ptrdiff_t i = 0; i < m.colsize(); ++i) s += double(m(i, j)) * double(m(i, j));
    return std::sqrt(s);
}

template <class T>
static double Norm(const tmv::Vector<T>& v)
{
    double s = 0;
    for (ptrdiff_t i = 0; i < v.size(); ++i) s += double(v(i)) * double(v(i));
    return std::sqrt(s);
}

template <class T, int S>
static tmv::Vector<T> MatVec(const tmv::Matrix<T, S>& m, const tmv::Vector<T>& x)
{
    tmv::Vector<T> y(m.colsize());
    for (ptrdiff_t i = 0; i < m.colsize(); ++i) {
        long double s = 0;
        for (ptrdiff_t j = 0; j < m.rowsize(); ++j) s += (long double)m(i, j) * x(j);
        y(i) = (T)s;
    }
    return y;
}

template <class T, int S>
static tmv::Vector<T> VecMat(const tmv::Vector<T>& x, const tmv::Matrix<T, S>& m)
{
    tmv::Vector<T> y(m.rowsize());
    for (ptrdiff_t j = 0; j < m.rowsize(); ++j) {
        long double s = 0;
        for (ptrdiff_t i = 0; i < m.colsize(); ++i) s += (long double)x(i) * m(i, j);
        y(j) = (T)s;
    }
    return y;
}

// tests/TMV_TestMatrixDiv.cpp:8-88, for both storage orders.
template <int S>
static void TestSquareDiv()
{
    tmv::Matrix<double, S> m(4, 4);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) m(i, j) = 2 + 4 * i - 5 * j;
    m(0, 0) = 14; m(1, 0) = -2; m(2, 0) = 7; m(3, 0) = -10; m(2, 2) = 30;
    tmv::Vector<double> b{2, -10, 5, -5};
    m.divideUsing(tmv::LU);
    m.saveDiv();
    m.setDiv();
    Assert(m.lud().isTrans() == (S == tmv::RowMajor), "isTrans follows the storage order");
    tmv::Vector<double> x = b / m;
    tmv::Vector<double> b2 = MatVec(m, x);
    tmv::Vector<double> d(4);
    for (int i = 0; i < 4; ++i) d(i) = b2(i) - b(i);
    Assert(Norm(d) < 1e3 * EPS * Norm(b), "Square b/m");
    x = b % m;
    b2 = VecMat(x, m);
    for (int i = 0; i < 4; ++i) d(i) = b2(i) - b(i);
    Assert(Norm(d) < 1e3 * EPS * Norm(b), "Square b%m");
    const double det = m.det();
    Assert(std::fabs(det - 28800) < 1e-9 * 28800, "Square Det = 28800");
    double sg;
    Assert(std::fabs(m.lud().logDet(&sg) - std::log(28800.0)) < 1e-12 && sg == 1.0,
           "Square logDet and sign");
}

// examples/Division.cpp:8-73, including the saveDiv behaviour.
static void TestDivisionExample()
{
    tmv::Matrix<double> A(4, 4);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) A(i, j) = 6. * i - 2 * j * j + 2.;
    for (int i = 0; i < 4; ++i) A(i, i) += 5.;
    tmv::Vector<double> b(4);
    for (int i = 0; i < 4; ++i) b(i) = 3. + 2. * i;
    auto close6 = [](const tmv::Vector<double>& x, std::initializer_list<double> g) {
        int i = 0;
        for (double v : g) { if (std::fabs(x(i) - v) > 5e-6 * std::fabs(v) + 1e-12) return false; ++i; }
        return true;
    };
    tmv::Vector<double> x = b / A;
    Assert(close6(x, {0.294545, 0.170909, 0.0472727, -0.0763636}), "Division: x = b/A");
    A(0, 0) = 20.;
    x = b / A;
    Assert(close6(x, {0.133458, 0.29877, 0.117091, -0.064587}), "Division: re-solve after change");
    tmv::Matrix<double> A2 = A;
    A2.saveDiv();
    x = b / A2;
    Assert(close6(x, {0.133458, 0.29877, 0.117091, -0.064587}), "Division: saveDiv x1");
    tmv::Vector<double> br{9, 7, 5, 3};
    x = br / A2;
    Assert(close6(x, {0.136753, 0.207381, -0.0775483, -0.362478}), "Division: reuse factors x2");
    for (int j = 0; j < 4; ++j) A2(0, j) *= 2.;
    x = b / A2;   // stale factorisation, as in TMV
    Assert(close6(x, {0.133458, 0.29877, 0.117091, -0.064587}), "Division: saveDiv keeps old factors");
    A2.resetDiv();
    x = b / A2;
    Assert(close6(x, {0.0703537, 0.348858, 0.144442, -0.0599736}), "Division: resetDiv refactorises");
}

// examples/Matrix.cpp:297-334 in every storage combination.
template <int SA, int SB>
static void TestProduct()
{
    const double m1v[12] = {42, 32, 33, 21, 26, 12, 33, 11, 8, 29, 27, 20};
    const double m5v[9] = {0, 1, 2, 2, 0, -2, 4, -1, -6};
    tmv::Matrix<double, SA> m1(4, 3);
    tmv::Matrix<double, SB> m5(3, 3);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 3; ++j) m1(i, j) = m1v[i * 3 + j];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m5(i, j) = m5v[i * 3 + j];
    tmv::Matrix<double> p = m1 * m5;
    const double g[12] = {196, 9, -178, 100, 9, -82, 54, 25, -4, 134, 9, -116};
    bool ok = true;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 3; ++j) ok = ok && p(i, j) == g[i * 3 + j];
    Assert(ok, "m1 * m5 (exact)");
    tmv::Matrix<double, tmv::RowMajor> c(4, 3, 1.0);
    tmv::MultMM<true>(-2.0, m1, m5, c);     // c += -2 m1 m5 into row-major C
    ok = true;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 3; ++j) ok = ok && c(i, j) == 1 - 2 * g[i * 3 + j];
    Assert(ok, "MultMM<true> into row-major C");
}

// tests/TMV_TestMatrixDet.cpp exact determinants.
static void TestDet()
{
    tmv::Matrix<double> m4(4, 4, {6, 1, 2, 4, 3, 2, 1, 8, 5, 6, 7, 1, 2, 9, 8, 3});
    Assert(std::fabs(m4.det() + 66) < 66 * 1e3 * EPS, "4x4 determinant -66");
    tmv::Matrix<double> m5(5, 5, {6, 1, 2, 4, 1, 3, 2, 1, 8, 4, 5, 6, 7, 1, 2, 2, 9, 8, 3, 7, 1, 4, 5, 9, 3});
    Assert(std::fabs(m5.det() + 1118) < 1118 * 1e3 * EPS, "5x5 determinant -1118");
}

// LU_Decompose + getP/getL/getU reconstruct PA = LU (TMV_TestMatrixDecomp).
static void TestDecomp()
{
    const int N = 200;
    tmv::Matrix<double> m(N, N);
    for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) m(i, j) = (2.0 + 4 * i - 5 * j) / (10 * N);
    m(3, 3) = 14; m(3, 4) = -2; m(0, 2) = 7; m(N - 1, N - 4) = 23; m(N - 1, N - 2) = 13; m(N - 1, N - 1) = -10;
    for (int i = 0; i < N; ++i) m(i, i) *= 30 * N;
    m.divideUsing(tmv::LU);
    m.saveDiv();
    const tmv::LUDiv<double>& lu = m.lud();
    tmv::Matrix<double> L = lu.getL(), U = lu.getU();
    tmv::Matrix<double> LUp = L * U;
    tmv::Matrix<double> PA = m;
    for (int i = 0; i < N; ++i) {
        const ptrdiff_t p = lu.getP()[i];
        if (p != i) for (int j = 0; j < N; ++j) std::swap(PA(i, j), PA(p, j));
    }
    double e = 0;
    for (int j = 0; j < N; ++j) for (int i = 0; i < N; ++i) e = std::fmax(e, std::fabs(PA(i, j) - LUp(i, j)));
    double amax = 0;
    for (int j = 0; j < N; ++j) for (int i = 0; i < N; ++i) amax = std::fmax(amax, std::fabs(m(i, j)));
    Assert(e <= 8 * N * std::numeric_limits<double>::epsilon() * amax, "P A = L U to 8 n eps");
    Assert(!lu.isSingular(), "regular fixture is not singular");
    tmv::Matrix<double> A2 = m;
    tmv::Permutation P;
    tmv::LU_Decompose(A2, P);
    bool same = true;
    for (int i = 0; i < N; ++i) same = same && P[i] == lu.getP()[i];
    Assert(same, "LU_Decompose(A, P) gives the divider's pivots");
}

static void TestSingular()
{
    tmv::Matrix<double> m(20, 20, 0.0);
    for (int i = 0; i < 20; ++i) for (int j = 0; j < 20; ++j) m(i, j) = 2.0 + 4 * i - 5 * j;
    for (int i = 0; i < 20; ++i) m(i, 1) = 0;
    tmv::Vector<double> b(20, 1.0);
    bool thrown = false;
    try { tmv::Vector<double> x = b / m; (void)x; }
    catch (tmv::Singular&) { thrown = true; }
    Assert(thrown, "b/m on a singular matrix throws tmv::Singular");
}

// examples/Division.cpp:75-83: A2.inverse() to the 6 printed digits, in
// both storage orders (a row-major matrix is factorised as its transpose);
// m.logDet(&sign) against log|det| of the same matrix; inverse of a
// singular matrix throws tmv::Singular.
template <int S>
static void TestInverse()
{
    tmv::Matrix<double, S> A2(4, 4);
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) A2(i, j) = 6. * i - 2 * j * j + 2.;
    for (int i = 0; i < 4; ++i) A2(i, i) += 5.;
    A2(0, 0) = 20.;
    for (int j = 0; j < 4; ++j) A2(0, j) *= 2.;
    const double g[16] = {0.0210347, -0.0395431, -0.012522, 0.0325132,
                          -0.016696, 0.0966608, -0.0527241, 0.0316344,
                          -0.00911687, -0.0643234, 0.139631, -0.0537786,
                          -0.00153779, -0.0253076, -0.0680141, 0.0608084};
    tmv::Matrix<double, S> X = A2.inverse();
    bool ok = true;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            ok = ok && std::fabs(X(i, j) - g[i * 4 + j]) <= 5e-6 * std::fabs(g[i * 4 + j]);
    Assert(ok, S == tmv::ColMajor ? "A2.inverse() (col-major)" : "A2.inverse() (row-major)");
    double sg = 0;
    const double ld = A2.logDet(&sg);
    Assert(std::fabs(sg * std::exp(ld) - A2.det()) <= 1e-12 * std::fabs(A2.det()) && sg != 0,
           "logDet(&sign) agrees with det()");
    tmv::Matrix<double, S> m(20, 20, 0.0);
    for (int i = 0; i < 20; ++i) for (int j = 0; j < 20; ++j) m(i, j) = 2.0 + 4 * i - 5 * j;
    for (int i = 0; i < 20; ++i) m(i, 1) = 0;
    bool thrown = false;
    try { tmv::Matrix<double, S> mi = m.inverse(); (void)mi; }
    catch (tmv::Singular&) { thrown = true; }
    Assert(thrown, "inverse of a singular matrix throws tmv::Singular");
    Assert(m.det() == 0.0 && m.logDet(&sg) == -std::numeric_limits<double>::infinity() && sg == 0,
           "singular: det 0, logDet -inf, sign 0");
}

// Complex products (TMV splits them into four real ones): integer parts,
// so the GPU result is exact, in both storage orders of C.
template <int SC>
static void TestComplexProduct()
{
    typedef std::complex<double> Z;
    tmv::Matrix<Z> a(3, 2), b(2, 2);
    a(0, 0) = Z(1, 2); a(0, 1) = Z(0, -1); a(1, 0) = Z(3, 0); a(1, 1) = Z(-2, 4);
    a(2, 0) = Z(5, -5); a(2, 1) = Z(1, 1);
    b(0, 0) = Z(2, 1); b(0, 1) = Z(-1, 0); b(1, 0) = Z(0, 3); b(1, 1) = Z(4, -2);
    tmv::Matrix<Z, SC> c(3, 2, Z(1, -1));
    tmv::MultMM<true>(Z(0, 1), a, b, c);                // c += i a b
    bool ok = true;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j) {
            Z s(0, 0);
            for (int l = 0; l < 2; ++l) s += a(i, l) * b(l, j);
            ok = ok && c(i, j) == Z(1, -1) + Z(0, 1) * s;
        }
    Assert(ok, SC == tmv::ColMajor ? "complex MultMM<true> (col-major C)"
                                   : "complex MultMM<true> (row-major C)");
}

// Complex LU division: x = b/m and x = b%m for a complex 3x3 with integer
// parts; A x reproduces b to a few eps.
static void TestComplexDivision()
{
    typedef std::complex<double> Z;
    tmv::Matrix<Z> m(3, 3);
    const Z v[9] = {Z(2, 1), Z(0, -3), Z(1, 1), Z(-1, 2), Z(4, 0), Z(0, 1), Z(3, -2), Z(1, 1), Z(-2, 5)};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m(i, j) = v[i * 3 + j];
    tmv::Vector<Z> b(3);
    b(0) = Z(1, 0); b(1) = Z(0, 2); b(2) = Z(-3, 1);
    tmv::Vector<Z> x = b / m;
    tmv::Vector<Z> y = b % m;
    double e1 = 0, e2 = 0;
    for (int i = 0; i < 3; ++i) {
        Z s1(0, 0), s2(0, 0);
        for (int j = 0; j < 3; ++j) { s1 += m(i, j) * x(j); s2 += y(j) * m(j, i); }
        e1 = std::fmax(e1, std::abs(s1 - b(i)));
        e2 = std::fmax(e2, std::abs(s2 - b(i)));
    }
    Assert(e1 < 1e3 * EPS && e2 < 1e3 * EPS, "complex b/m and b%m");
}

// QR division: the default for a non-square matrix (least squares), and on
// request for a square one (same x as LU).
static void TestQRDivision()
{
    tmv::Matrix<double> a(4, 2, {1, 0, 0, 1, 1, 1, 1, -1});
    tmv::Vector<double> b{1, 2, 3, 4};
    tmv::Vector<double> x = b / a;        // A^T A = 3 I, A^T b = (8, 1)
    Assert(std::fabs(x(0) - 8.0 / 3) < 1e3 * EPS && std::fabs(x(1) - 1.0 / 3) < 1e3 * EPS,
           "QR least squares (default for non-square)");
    tmv::Matrix<double> m(3, 3, {4, 1, 2, 1, 5, 3, 2, 3, 6});
    tmv::Vector<double> c{1, -2, 3};
    const tmv::Vector<double> xl = c / m;
    m.divideUsing(tmv::QR);
    const tmv::Vector<double> xq = c / m;
    double d = 0;
    for (int i = 0; i < 3; ++i) d = std::fmax(d, std::fabs(xl(i) - xq(i)));
    Assert(d < 1e3 * EPS, "QR division of a square matrix agrees with LU");
}

static void TestMixed()
{
    const int N = 300;
    tmv::Matrix<double> m(N, N);
    for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) m(i, j) = std::sin(1.0 + i * 0.37 + j * j * 0.011);
    for (int i = 0; i < N; ++i) m(i, i) += N;
    tmv::Vector<double> b(N);
    for (int i = 0; i < N; ++i) b(i) = std::cos(0.1 * i);
    tmvx_ir_info info;
    tmv::Vector<double> x = tmv::SolveMixed(b, m, 1e-12, &info);
    tmv::Vector<double> x64 = b / m;
    double d = 0, xm = 0;
    for (int i = 0; i < N; ++i) { d = std::fmax(d, std::fabs(x(i) - x64(i))); xm = std::fmax(xm, std::fabs(x64(i))); }
    Assert(info.converged && info.backward_error <= 1e-12 && d <= 1e-12 * xm,
           "SolveMixed agrees with b/m");
}

// examples/Matrix.cpp:326-342: m5 *= m5 is m5 = m5 * m5 although C aliases
// both operands (TMV copies; the GPU path stages the operands first).
template <int S>
static void TestAliasProduct()
{
    const double m5v[9] = {0, 1, 2, 2, 0, -2, 4, -1, -6};
    const double g[9] = {10, -2, -14, -8, 4, 16, -26, 10, 46};
    tmv::Matrix<double, S> m5(3, 3);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m5(i, j) = m5v[i * 3 + j];
    tmv::Matrix<double> p = m5 * m5;
    m5 *= m5;
    bool ok = true;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ok = ok && m5(i, j) == g[i * 3 + j] && p(i, j) == g[i * 3 + j];
    Assert(ok, S == tmv::ColMajor ? "m5 *= m5 (col-major)" : "m5 *= m5 (row-major)");
    m5 *= 0.5;
    ok = true;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) ok = ok && m5(i, j) == 0.5 * g[i * 3 + j];
    Assert(ok, "m *= x");
}

// Integer-valued operands: every product below is exact in double, so the
// GPU result must equal the host one bit for bit.
static double ival(int i, int j, int salt) { return (double)(((i * 7 + j * 13 + salt * 5) % 17) - 8); }

template <class T, class MA, class MB>
static tmv::Matrix<T> HostProd(T x, const MA& A, const MB& B)
{
    tmv::Matrix<T> C(A.colsize(), B.rowsize());
    for (ptrdiff_t i = 0; i < C.colsize(); ++i)
        for (ptrdiff_t j = 0; j < C.rowsize(); ++j) {
            T s(0);
            for (ptrdiff_t l = 0; l < A.rowsize(); ++l) s += T(A(i, l)) * T(B(l, j));
            C(i, j) = x * s;
        }
    return C;
}

template <class T, class MC>
static bool Same(const tmv::Matrix<T>& R, const MC& C)
{
    if (R.colsize() != C.colsize() || R.rowsize() != C.rowsize()) return false;
    for (ptrdiff_t i = 0; i < R.colsize(); ++i)
        for (ptrdiff_t j = 0; j < R.rowsize(); ++j) if (!(T(C(i, j)) == R(i, j))) return false;
    return true;
}

// C = A*B, x*A*B, A*B*x, C += A*B, C -= A*B, C += x*A*B, (A*B)*D
// (TMV_MatrixArith.h:700-812) into either storage order.
template <int SC>
static void TestProductForms()
{
    const int M = 37, K = 23, N = 29;
    tmv::Matrix<double> A(M, K), D(N, 11);
    tmv::Matrix<double, tmv::RowMajor> B(K, N);
    tmv::Matrix<double, SC> C0(M, N);
    for (int i = 0; i < M; ++i) for (int j = 0; j < K; ++j) A(i, j) = ival(i, j, 1);
    for (int i = 0; i < K; ++i) for (int j = 0; j < N; ++j) B(i, j) = ival(i, j, 2);
    for (int i = 0; i < M; ++i) for (int j = 0; j < N; ++j) C0(i, j) = ival(i, j, 3);
    for (int i = 0; i < N; ++i) for (int j = 0; j < 11; ++j) D(i, j) = ival(i, j, 4);
    const tmv::Matrix<double> AB = HostProd(1.0, A, B);
    tmv::Matrix<double, SC> C(M, N);
    C = A * B;
    Assert(Same(AB, C), "C = A*B");
    C = 2.5 * A * B;
    Assert(Same(HostProd(2.5, A, B), C), "C = x*A*B");
    C = A * B * -3.0;
    Assert(Same(HostProd(-3.0, A, B), C), "C = A*B*x");
    tmv::Matrix<double> R(M, N);
    C = C0;
    C += A * B;
    for (int i = 0; i < M; ++i) for (int j = 0; j < N; ++j) R(i, j) = C0(i, j) + AB(i, j);
    Assert(Same(R, C), "C += A*B");
    C = C0;
    C -= A * B;
    for (int i = 0; i < M; ++i) for (int j = 0; j < N; ++j) R(i, j) = C0(i, j) - AB(i, j);
    Assert(Same(R, C), "C -= A*B");
    C = C0;
    C += 0.25 * A * B;
    for (int i = 0; i < M; ++i) for (int j = 0; j < N; ++j) R(i, j) = C0(i, j) + 0.25 * AB(i, j);
    Assert(Same(R, C), "C += x*A*B");
    tmv::Matrix<double> E = (A * B) * D;
    Assert(Same(HostProd(1.0, AB, D), E), "(A*B)*D");
}

// Views: sub-blocks (unit stride, leading dimension larger than the
// block), strided blocks (no unit stride: packed on the way), transposes
// and row-major blocks, as operands and as the target; the target's
// neighbours stay untouched.
static void TestViewProducts()
{
    tmv::Matrix<double> big(60, 70);
    tmv::Matrix<double, tmv::RowMajor> bigr(50, 45);
    for (int i = 0; i < 60; ++i) for (int j = 0; j < 70; ++j) big(i, j) = ival(i, j, 5);
    for (int i = 0; i < 50; ++i) for (int j = 0; j < 45; ++j) bigr(i, j) = ival(i, j, 6);
    const tmv::Matrix<double>& cbig = big;
    tmv::ConstMatrixView<double> A = cbig.subMatrix(3, 30, 5, 24);            // 27 x 19, cm
    tmv::ConstMatrixView<double> At = cbig.subMatrix(2, 21, 40, 67).transpose(); // 27 x 19, rm
    tmv::ConstMatrixView<double> As = cbig.subMatrix(1, 55, 0, 57, 2, 3);      // 27 x 19, strided
    tmv::ConstMatrixView<double> B = bigr.subMatrix(4, 23, 10, 41);            // 19 x 31, rm
    tmv::ConstMatrixView<double> Bs = bigr.subMatrix(0, 38, 1, 32, 2, 1).transpose().transpose();
    struct Case { const char* name; tmv::ConstMatrixView<double> a, b; };
    const Case cases[] = {{"view (cm) * view (rm)", A, B}, {"transposed view * view", At, B},
                          {"strided view * view", As, B}, {"view * strided view", A, Bs},
                          {"strided * strided", As, Bs}};
    for (const Case& c : cases) {
        tmv::Matrix<double> C = c.a * c.b;
        Assert(Same(HostProd(1.0, c.a, c.b), C), c.name);
    }
    // target: a block of a bigger matrix, a strided block, a transposed block
    tmv::Matrix<double> tgt(80, 90, 7.0);
    tmv::Matrix<double> before = tgt;
    tgt.subMatrix(10, 37, 20, 51) += A * B;
    tmv::Matrix<double> R = before;
    const tmv::Matrix<double> AB = HostProd(1.0, A, B);
    for (int i = 0; i < 27; ++i) for (int j = 0; j < 31; ++j) R(10 + i, 20 + j) += AB(i, j);
    Assert(Same(R, tgt), "block target += A*B, neighbours untouched");
    tgt = before;
    tgt.subMatrix(0, 54, 0, 62, 2, 2) = As * B;
    R = before;
    const tmv::Matrix<double> AsB = HostProd(1.0, As, B);
    for (int i = 0; i < 27; ++i) for (int j = 0; j < 31; ++j) R(2 * i, 2 * j) = AsB(i, j);
    Assert(Same(R, tgt), "strided target = A*B");
    tgt = before;
    tgt.subMatrix(5, 36, 7, 34).transpose() -= 2.0 * A * B;
    R = before;
    for (int i = 0; i < 27; ++i) for (int j = 0; j < 31; ++j) R(5 + j, 7 + i) -= 2.0 * AB(i, j);
    Assert(Same(R, tgt), "transposed block target -= x*A*B");
    // MultMM<add> on views, C aliasing an operand block
    tmv::Matrix<double> sq(30, 30);
    for (int i = 0; i < 30; ++i) for (int j = 0; j < 30; ++j) sq(i, j) = ival(i, j, 7);
    const tmv::Matrix<double> sq0 = sq;
    tmv::MultMM<true>(1.0, sq.view(), sq.transpose(), sq.view());
    R = HostProd(1.0, sq0, sq0.transpose());
    for (int i = 0; i < 30; ++i) for (int j = 0; j < 30; ++j) R(i, j) += sq0(i, j);
    Assert(Same(R, sq), "MultMM<true>(1, m, m^T, m): C aliases both operands");
}

// Conjugated and adjoint views of complex matrices, as operands and target.
static void TestConjViews()
{
    typedef std::complex<double> Z;
    tmv::Matrix<Z> A(13, 9), B(9, 11), C0(13, 11);
    for (int i = 0; i < 13; ++i) for (int j = 0; j < 9; ++j) A(i, j) = Z(ival(i, j, 1), ival(i, j, 2));
    for (int i = 0; i < 9; ++i) for (int j = 0; j < 11; ++j) B(i, j) = Z(ival(i, j, 3), ival(i, j, 4));
    for (int i = 0; i < 13; ++i) for (int j = 0; j < 11; ++j) C0(i, j) = Z(ival(i, j, 5), ival(i, j, 6));
    tmv::Matrix<Z> Ac = A.conjugate();                 // explicit conj(A) copy
    tmv::Matrix<Z> C = A.conjugate() * B;
    Assert(Same(HostProd(Z(1), Ac, B), C), "conj(A) * B");
    tmv::Matrix<Z> Ah(9, 13);
    for (int i = 0; i < 9; ++i) for (int j = 0; j < 13; ++j) Ah(i, j) = std::conj(A(j, i));
    tmv::Matrix<Z> G = A.adjoint() * C0;
    Assert(Same(HostProd(Z(1), Ah, C0), G), "A^H * C");
    tmv::Matrix<Z> T = C0;
    T.conjugate() += Z(0, 1) * A * B;                  // conj(T) += i A B
    const tmv::Matrix<Z> AB = HostProd(Z(1), A, B);
    tmv::Matrix<Z> R(13, 11);
    for (int i = 0; i < 13; ++i)
        for (int j = 0; j < 11; ++j) R(i, j) = std::conj(std::conj(C0(i, j)) + Z(0, 1) * AB(i, j));
    Assert(Same(R, T), "conj(C) += x A B");
}

// LU_Decompose on views (src/TMV_LUDecompose.cpp:514-538): a row-major
// matrix, a block of a bigger matrix and a strided block give the factors
// and pivots of the matrix they show (the same GPU factorisation as a
// dense column-major copy, so bit-identical); a conjugated view gets the
// conjugated factors with the same pivots.
static void TestLUViews()
{
    const int N = 70;
    tmv::Matrix<double> Mc(N, N);
    for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) Mc(i, j) = std::sin(0.3 + 0.71 * i + 0.37 * j * j);
    tmv::Matrix<double> LUc = Mc;
    tmv::Permutation Pc;
    tmv::LU_Decompose(LUc, Pc);
    auto samep = [&](const tmv::Permutation& P) {
        for (int i = 0; i < N; ++i) if (P[i] != Pc[i]) return false;
        return true;
    };
    tmv::Matrix<double, tmv::RowMajor> Mr(Mc);
    tmv::Permutation Pr;
    tmv::LU_Decompose(Mr, Pr);
    Assert(samep(Pr) && Same(LUc, Mr), "LU_Decompose of a row-major matrix");
    tmv::Matrix<double> big(N + 9, N + 13, 3.0);
    big.subMatrix(4, 4 + N, 6, 6 + N) = Mc.view();
    const tmv::Matrix<double> big0 = big;
    tmv::Permutation Pb;
    tmv::LU_Decompose(big.subMatrix(4, 4 + N, 6, 6 + N), Pb);
    bool ok = samep(Pb) && Same(LUc, big.subMatrix(4, 4 + N, 6, 6 + N));
    for (int i = 0; i < N + 9; ++i)
        for (int j = 0; j < N + 13; ++j)
            if (i < 4 || i >= 4 + N || j < 6 || j >= 6 + N) ok = ok && big(i, j) == big0(i, j);
    Assert(ok, "LU_Decompose of a block (ld > n), neighbours untouched");
    tmv::Matrix<double> sbig(2 * N, 3 * N, 0.0);
    sbig.subMatrix(0, 2 * N, 0, 3 * N, 2, 3) = Mc.view();
    tmv::Permutation Ps;
    tmv::LU_Decompose(sbig.subMatrix(0, 2 * N, 0, 3 * N, 2, 3), Ps);
    Assert(samep(Ps) && Same(LUc, sbig.subMatrix(0, 2 * N, 0, 3 * N, 2, 3)),
           "LU_Decompose of a strided view");
    typedef std::complex<double> Z;
    tmv::Matrix<Z> Z0(40, 40);
    for (int i = 0; i < 40; ++i)
        for (int j = 0; j < 40; ++j) Z0(i, j) = Z(std::sin(1.0 + i + 0.5 * j * j), std::cos(0.3 * i * j + 0.2));
    tmv::Matrix<Z> Zl = Z0;
    tmv::Permutation Pz, Pzc;
    tmv::LU_Decompose(Zl, Pz);
    tmv::Matrix<Z> Zc = Z0;
    tmv::LU_Decompose(Zc.conjugate(), Pzc);
    ok = true;
    for (int i = 0; i < 40; ++i) {
        ok = ok && Pz[i] == Pzc[i];
        for (int j = 0; j < 40; ++j) ok = ok && Zc(i, j) == Zl(i, j);   // storage: LU of Z0
    }
    Assert(ok, "LU_Decompose of a conjugated view: conj(A)'s factors are conj(A's)");
}

// LU_Decompose on non-square matrices and views: P A = L U with L m x r
// unit-lower and U r x n upper (r = min(m, n)), in both storage orders.
template <int S>
static void TestLURect(int M, int N)
{
    tmv::Matrix<double, S> A(M, N);
    for (int i = 0; i < M; ++i) for (int j = 0; j < N; ++j) A(i, j) = std::sin(0.7 + 1.3 * i + 0.29 * j * j);
    tmv::Matrix<double, S> LU = A;
    tmv::Permutation P;
    tmv::LU_Decompose(LU, P);
    const int r = M < N ? M : N;
    tmv::Matrix<double> PA(A);
    for (int i = 0; i < r; ++i) {
        const ptrdiff_t p = P[i];
        if (p != i) for (int j = 0; j < N; ++j) std::swap(PA(i, j), PA(p, j));
    }
    double e = 0, amax = 0;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            long double s = 0;
            for (int k = 0; k <= std::min(i, std::min(j, r - 1)); ++k)
                s += (long double)(k == i ? 1.0 : LU(i, k)) * (k <= j ? LU(k, j) : 0.0);
            e = std::fmax(e, std::fabs((double)(PA(i, j) - s)));
            amax = std::fmax(amax, std::fabs(A(i, j)));
        }
    bool identity_tail = P.size() == M;
    for (int i = r; i < M; ++i) identity_tail = identity_tail && P[i] == i;
    char name[96];
    std::snprintf(name, sizeof(name), "LU_Decompose %dx%d (%s): P A = L U", M, N,
                  S == tmv::ColMajor ? "col-major" : "row-major");
    Assert(e <= 8 * std::max(M, N) * std::numeric_limits<double>::epsilon() * amax && identity_tail,
           name);
}

// LUDiv::isSingular on the device and LUDiv::checkDecomp (src/TMV_LUD.cpp:
// 195-254), in both storage orders and for complex.
template <int S>
static void TestLUDivChecks()
{
    const int N = 64;
    tmv::Matrix<double, S> m(N, N);
    for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) m(i, j) = std::cos(0.2 + 0.9 * i * j + 0.1 * j);
    for (int i = 0; i < N; ++i) m(i, i) += 3.0;
    m.divideUsing(tmv::LU);
    m.saveDiv();
    Assert(!m.lud().isSingular(), "isSingular: regular matrix");
    Assert(m.lud().checkDecomp(m), S == tmv::ColMajor ? "checkDecomp (col-major)"
                                                      : "checkDecomp (row-major, isTrans)");
    tmv::Matrix<double, S> m2 = m;
    m2(5, 7) += 1.0;
    Assert(!m.lud().checkDecomp(m2), "checkDecomp rejects a different matrix");
    tmv::Matrix<double, S> s(N, N);
    for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) s(i, j) = 1.0 + i + 2.0 * j;   // rank 2
    tmv::LUDiv<double> sd(s);
    Assert(sd.isSingular(), "isSingular: rank-2 matrix");
    typedef std::complex<double> Z;
    tmv::Matrix<Z> z(20, 20);
    for (int i = 0; i < 20; ++i) for (int j = 0; j < 20; ++j) z(i, j) = Z(std::sin(i + 0.3 * j), std::cos(0.7 * i * j));
    for (int i = 0; i < 20; ++i) z(i, i) += Z(4, 1);
    tmv::LUDiv<Z> zd(z);
    Assert(!zd.isSingular() && zd.checkDecomp(z), "complex isSingular / checkDecomp");
    for (int i = 0; i < 20; ++i) z(i, 3) = Z(0, 0);
    tmv::LUDiv<Z> zs(z);
    Assert(zs.isSingular(), "complex isSingular: zero column");
    // LUDiv from a view (a block of a bigger matrix) solves like the copy
    tmv::Matrix<double> big(N + 5, N + 5, 1.0);
    big.subMatrix(2, 2 + N, 3, 3 + N) = m.view();
    tmv::LUDiv<double> bd(big.subMatrix(2, 2 + N, 3, 3 + N));
    tmv::Vector<double> b(N), x1(N);
    for (int i = 0; i < N; ++i) b(i) = std::sin(0.5 * i);
    x1 = b / m;
    tmv::Vector<double> x2 = b;
    bd.lDivEqCM(x2.ptr(), 1);
    double d = 0;
    for (int i = 0; i < N; ++i) d = std::fmax(d, std::fabs(x1(i) - x2(i)));
    Assert(d <= 1e-12, "LUDiv of a view solves as the matrix does");
}

int main()
{
    TestSquareDiv<tmv::ColMajor>();
    TestSquareDiv<tmv::RowMajor>();
    TestDivisionExample();
    TestProduct<tmv::ColMajor, tmv::ColMajor>();
    TestProduct<tmv::ColMajor, tmv::RowMajor>();
    TestProduct<tmv::RowMajor, tmv::ColMajor>();
    TestProduct<tmv::RowMajor, tmv::RowMajor>();
    TestDet();
    TestDecomp();
    TestSingular();
    TestMixed();
    TestInverse<tmv::ColMajor>();
    TestInverse<tmv::RowMajor>();
    TestComplexProduct<tmv::ColMajor>();
    TestComplexProduct<tmv::RowMajor>();
    TestComplexDivision();
    TestQRDivision();
    TestAliasProduct<tmv::ColMajor>();
    TestAliasProduct<tmv::RowMajor>();
    TestProductForms<tmv::ColMajor>();
    TestProductForms<tmv::RowMajor>();
    TestViewProducts();
    TestConjViews();
    TestLUViews();
    TestLURect<tmv::ColMajor>(150, 70);
    TestLURect<tmv::ColMajor>(70, 150);
    TestLURect<tmv::RowMajor>(150, 70);
    TestLURect<tmv::RowMajor>(70, 150);
    TestLUDivChecks<tmv::ColMajor>();
    TestLUDivChecks<tmv::RowMajor>();
    std::printf("%s: %d failure(s)\n", nfail ? "FAILED" : "PASSED", nfail);
    return nfail ? 1 : 0;
}
