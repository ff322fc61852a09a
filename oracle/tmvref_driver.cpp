// Thin extern "C" driver over the real TMV 0.76 library (built from the
// reference sources by oracle/Makefile.ref into oracle/_ref/libtmv_ref.so).
//
// TEST INFRASTRUCTURE ONLY.  Tests use it to pin the oracle restatement
// (oracle/tmv_oracle.cpp) against TMV itself, and bench.py's CPU-baseline leg
// times it.  Nothing in the product links or loads it.
//
// Every entry point views the caller's buffers with TMV's own MatrixView
// (ptr, colsize, rowsize, stepi, stepj, ct) -- TMV_Matrix.h:1335-1345 -- and
// calls the TMV routine named in its comment.  Exceptions never cross this
// boundary: they become negative return codes.
#include "TMV.h"
#include "TMV_LUDiv.h"   // LU_LDivEq / LU_RDivEq (src/TMV_LUDiv.h:29-35)
#include "TMV_QRDiv.h"   // QR_Decompose(A, beta, det) (src/TMV_QRDiv.h:34-41)
#include <cstring>
#include <omp.h>

namespace {

template <class T>
tmv::ConstMatrixView<T> cview(const T* p, ptrdiff_t rows, ptrdiff_t cols,
                              ptrdiff_t ld, int rowmajor)
{
    return rowmajor ? tmv::ConstMatrixView<T>(p, rows, cols, ld, 1, tmv::NonConj)
                    : tmv::ConstMatrixView<T>(p, rows, cols, 1, ld, tmv::NonConj);
}

template <class T>
tmv::MatrixView<T> mview(T* p, ptrdiff_t rows, ptrdiff_t cols, ptrdiff_t ld,
                         int rowmajor)
{
    return rowmajor ? tmv::MatrixView<T>(p, rows, cols, ld, 1, tmv::NonConj)
                    : tmv::MatrixView<T>(p, rows, cols, 1, ld, tmv::NonConj);
}

// C (+)= alpha*A*B through tmv::MultMM<add> (src/TMV_MultMM.cpp:404-477).
template <class T>
int gemm(int a_rm, int b_rm, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k, T alpha,
         const T* A, ptrdiff_t lda, const T* B, ptrdiff_t ldb, int add, T* C,
         ptrdiff_t ldc)
{
    try {
        tmv::ConstMatrixView<T> Av = cview(A, m, k, lda, a_rm);
        tmv::ConstMatrixView<T> Bv = cview(B, k, n, ldb, b_rm);
        tmv::MatrixView<T> Cv = mview(C, m, n, ldc, 0);
        if (add) tmv::MultMM<true>(alpha, Av, Bv, Cv);
        else tmv::MultMM<false>(alpha, Av, Bv, Cv);
    } catch (...) { return -1; }
    return 0;
}

// Complex C (+)= alpha*op(A)*op(B) through tmv::MultMM<add> on
// std::complex views, op = conjugate() when conj_x (TMV's complex block split,
// src/TMV_MultMM_Block.cpp:334-392).
template <class RT>
int gemm_cplx(int a_rm, int b_rm, int conj_a, int conj_b, ptrdiff_t m, ptrdiff_t n,
              ptrdiff_t k, RT are, RT aim, const RT* A, ptrdiff_t lda, const RT* B,
              ptrdiff_t ldb, int add, RT* C, ptrdiff_t ldc)
{
    typedef std::complex<RT> CT;
    try {
        const tmv::ConstMatrixView<CT> A0 = cview(reinterpret_cast<const CT*>(A), m, k, lda, a_rm);
        const tmv::ConstMatrixView<CT> B0 = cview(reinterpret_cast<const CT*>(B), k, n, ldb, b_rm);
        const tmv::ConstMatrixView<CT> Av = conj_a ? A0.conjugate() : A0;
        const tmv::ConstMatrixView<CT> Bv = conj_b ? B0.conjugate() : B0;
        tmv::MatrixView<CT> Cv = mview(reinterpret_cast<CT*>(C), m, n, ldc, 0);
        if (add) tmv::MultMM<true>(CT(are, aim), Av, Bv, Cv);
        else tmv::MultMM<false>(CT(are, aim), Av, Bv, Cv);
    } catch (...) { return -1; }
    return 0;
}

// QR through tmv::QR_Decompose(A, beta, det) (src/TMV_QRDecompose.cpp:603-650).
template <class T>
int geqrf(ptrdiff_t m, ptrdiff_t n, T* A, ptrdiff_t lda, T* beta)
{
    try {
        T det(1);
        tmv::QR_Decompose(mview(A, m, n, lda, 0), tmv::VectorView<T>(beta, n, 1, tmv::NonConj), det);
    } catch (...) { return -1; }
    return 0;
}

// x = b/A with QR division (least squares for m > n), src/TMV_QRDiv.cpp:33-94.
template <class T>
int qr_solve(ptrdiff_t m, ptrdiff_t n, ptrdiff_t nrhs, const T* A, ptrdiff_t lda, const T* B,
             ptrdiff_t ldb, T* X, ptrdiff_t ldx)
{
    try {
        tmv::Matrix<T, tmv::ColMajor> Am(m, n);
        Am = cview(A, m, n, lda, 0);
        Am.divideUsing(tmv::QR);
        tmv::Matrix<T, tmv::ColMajor> Bm(m, nrhs);
        Bm = cview(B, m, nrhs, ldb, 0);
        tmv::Matrix<T, tmv::ColMajor> Xm = Bm / Am;
        mview(X, n, nrhs, ldx, 0) = Xm;
    } catch (tmv::Singular&) { return -2;
    } catch (...) { return -1; }
    return 0;
}

// In-place PA=LU through tmv::LU_Decompose (src/TMV_LUDecompose.cpp:514-559).
template <class T>
int getrf(ptrdiff_t n, T* A, ptrdiff_t lda, ptrdiff_t* P)
{
    try {
        tmv::LU_Decompose(mview(A, n, n, lda, 0), P);
    } catch (...) { return -1; }
    return 0;
}

// The same on an m x n view (R = min(m, n) pivots).
template <class T>
int getrf_mn(ptrdiff_t m, ptrdiff_t n, T* A, ptrdiff_t lda, ptrdiff_t* P)
{
    try {
        tmv::LU_Decompose(mview(A, m, n, lda, 0), P);
    } catch (...) { return -1; }
    return 0;
}

// B := (P L U)^-1 B through tmv::LU_LDivEq (src/TMV_LUDiv.cpp:185-226);
// trans != 0 solves with the transpose through LU_RDivEq on B^T
// (src/TMV_LUDiv.cpp:369-410): x^T (P L U) = b^T.
template <class T>
int getrs(int trans, ptrdiff_t n, ptrdiff_t nrhs, const T* LU, ptrdiff_t lda,
          const ptrdiff_t* P, T* B, ptrdiff_t ldb)
{
    try {
        tmv::ConstMatrixView<T> LUv = cview(LU, n, n, lda, 0);
        if (!trans) tmv::LU_LDivEq(LUv, P, mview(B, n, nrhs, ldb, 0));
        else tmv::LU_RDivEq(LUv, P, mview(B, n, nrhs, ldb, 0).transpose());
    } catch (tmv::Singular&) { return -2;
    } catch (...) { return -1; }
    return 0;
}

// x = b/A (nrhs == 1, the Vector path, TMV_MatrixArith.h:979-1012) or
// X = B/A (TMV_MatrixArith.h:1144-1177) with A.divideUsing(tmv::LU): the
// complete user-level call, factorisation included (src/TMV_Matrix.cpp:66-96).
template <class T>
int gesv(ptrdiff_t n, ptrdiff_t nrhs, const T* A, ptrdiff_t lda, const T* B,
         ptrdiff_t ldb, T* X, ptrdiff_t ldx)
{
    try {
        tmv::Matrix<T, tmv::ColMajor> Am(n, n);
        Am = cview(A, n, n, lda, 0);
        Am.divideUsing(tmv::LU);
        if (nrhs == 1) {
            tmv::Vector<T> b(n);
            for (ptrdiff_t i = 0; i < n; ++i) b(i) = B[i];
            tmv::Vector<T> x = b / Am;
            for (ptrdiff_t i = 0; i < n; ++i) X[i] = x(i);
        } else {
            tmv::Matrix<T, tmv::ColMajor> Bm(n, nrhs);
            Bm = cview(B, n, nrhs, ldb, 0);
            tmv::Matrix<T, tmv::ColMajor> Xm = Bm / Am;
            mview(X, n, nrhs, ldx, 0) = Xm;
        }
    } catch (tmv::Singular&) { return -2;
    } catch (...) { return -1; }
    return 0;
}

// det(A) through A.det() with LU division (src/TMV_LUD.cpp:132-156).
template <class T>
int det(ptrdiff_t n, const T* A, ptrdiff_t lda, T* out)
{
    try {
        tmv::Matrix<T, tmv::ColMajor> Am(n, n);
        Am = cview(A, n, n, lda, 0);
        Am.divideUsing(tmv::LU);
        *out = Am.det();
    } catch (...) { return -1; }
    return 0;
}

// A^-1 through A.inverse() with LU division (src/TMV_LUD.cpp:158-169,
// src/TMV_LUInverse.cpp:235-275).
template <class T>
int inverse(ptrdiff_t n, const T* A, ptrdiff_t lda, T* X, ptrdiff_t ldx)
{
    try {
        tmv::Matrix<T, tmv::ColMajor> Am(n, n);
        Am = cview(A, n, n, lda, 0);
        Am.divideUsing(tmv::LU);
        tmv::Matrix<T, tmv::ColMajor> Xm = Am.inverse();
        mview(X, n, n, ldx, 0) = Xm;
    } catch (tmv::Singular&) { return -2;
    } catch (...) { return -1; }
    return 0;
}

// log|det A| and its sign through A.logDet(&sign) (src/TMV_LUD.cpp:145-156).
template <class T>
int logdet(ptrdiff_t n, const T* A, ptrdiff_t lda, T* out, T* sign)
{
    try {
        tmv::Matrix<T, tmv::ColMajor> Am(n, n);
        Am = cview(A, n, n, lda, 0);
        Am.divideUsing(tmv::LU);
        *out = Am.logDet(sign);
    } catch (...) { return -1; }
    return 0;
}

} // namespace

extern "C" {

int tmvref_num_threads(void) { return omp_get_max_threads(); }
void tmvref_set_threads(int n) { omp_set_num_threads(n); }

int tmvref_gemm_f64(int a_rm, int b_rm, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                    double alpha, const double* A, ptrdiff_t lda,
                    const double* B, ptrdiff_t ldb, int add, double* C,
                    ptrdiff_t ldc)
{ return gemm<double>(a_rm, b_rm, m, n, k, alpha, A, lda, B, ldb, add, C, ldc); }

int tmvref_gemm_f32(int a_rm, int b_rm, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                    float alpha, const float* A, ptrdiff_t lda, const float* B,
                    ptrdiff_t ldb, int add, float* C, ptrdiff_t ldc)
{ return gemm<float>(a_rm, b_rm, m, n, k, alpha, A, lda, B, ldb, add, C, ldc); }

int tmvref_gemm_z(int a_rm, int b_rm, int conj_a, int conj_b, ptrdiff_t m, ptrdiff_t n,
                  ptrdiff_t k, double are, double aim, const double* A, ptrdiff_t lda,
                  const double* B, ptrdiff_t ldb, int add, double* C, ptrdiff_t ldc)
{ return gemm_cplx<double>(a_rm, b_rm, conj_a, conj_b, m, n, k, are, aim, A, lda, B, ldb, add, C, ldc); }

int tmvref_getrf_z(ptrdiff_t n, double* A, ptrdiff_t lda, ptrdiff_t* P)
{ return getrf<std::complex<double>>(n, reinterpret_cast<std::complex<double>*>(A), lda, P); }

int tmvref_geqrf_f64(ptrdiff_t m, ptrdiff_t n, double* A, ptrdiff_t lda, double* beta)
{ return geqrf<double>(m, n, A, lda, beta); }
int tmvref_qr_solve_f64(ptrdiff_t m, ptrdiff_t n, ptrdiff_t nrhs, const double* A, ptrdiff_t lda,
                        const double* B, ptrdiff_t ldb, double* X, ptrdiff_t ldx)
{ return qr_solve<double>(m, n, nrhs, A, lda, B, ldb, X, ldx); }

int tmvref_getrf_f64(ptrdiff_t n, double* A, ptrdiff_t lda, ptrdiff_t* P)
{ return getrf<double>(n, A, lda, P); }
int tmvref_getrf_f32(ptrdiff_t n, float* A, ptrdiff_t lda, ptrdiff_t* P)
{ return getrf<float>(n, A, lda, P); }
int tmvref_getrf_mn_f64(ptrdiff_t m, ptrdiff_t n, double* A, ptrdiff_t lda, ptrdiff_t* P)
{ return getrf_mn<double>(m, n, A, lda, P); }
int tmvref_getrf_mn_f32(ptrdiff_t m, ptrdiff_t n, float* A, ptrdiff_t lda, ptrdiff_t* P)
{ return getrf_mn<float>(m, n, A, lda, P); }

int tmvref_getrs_f64(int trans, ptrdiff_t n, ptrdiff_t nrhs, const double* LU,
                     ptrdiff_t lda, const ptrdiff_t* P, double* B, ptrdiff_t ldb)
{ return getrs<double>(trans, n, nrhs, LU, lda, P, B, ldb); }
int tmvref_getrs_f32(int trans, ptrdiff_t n, ptrdiff_t nrhs, const float* LU,
                     ptrdiff_t lda, const ptrdiff_t* P, float* B, ptrdiff_t ldb)
{ return getrs<float>(trans, n, nrhs, LU, lda, P, B, ldb); }

int tmvref_gesv_f64(ptrdiff_t n, ptrdiff_t nrhs, const double* A, ptrdiff_t lda,
                    const double* B, ptrdiff_t ldb, double* X, ptrdiff_t ldx)
{ return gesv<double>(n, nrhs, A, lda, B, ldb, X, ldx); }
int tmvref_gesv_f32(ptrdiff_t n, ptrdiff_t nrhs, const float* A, ptrdiff_t lda,
                    const float* B, ptrdiff_t ldb, float* X, ptrdiff_t ldx)
{ return gesv<float>(n, nrhs, A, lda, B, ldb, X, ldx); }

int tmvref_det_f64(ptrdiff_t n, const double* A, ptrdiff_t lda, double* out)
{ return det<double>(n, A, lda, out); }

int tmvref_inverse_f64(ptrdiff_t n, const double* A, ptrdiff_t lda, double* X, ptrdiff_t ldx)
{ return inverse<double>(n, A, lda, X, ldx); }

int tmvref_logdet_f64(ptrdiff_t n, const double* A, ptrdiff_t lda, double* out, double* sign)
{ return logdet<double>(n, A, lda, out, sign); }

} // extern "C"
