// tmvx_tmv.h -- TMV-shaped C++ API over the tmvx C ABI (header only).
//
// The part of TMV 0.76's user surface that the dense GEMM + LU path serves,
// with TMV's names, argument meaning and error behaviour, running on the
// MI355X through include/tmvx.h:
//   tmv::Matrix<T, tmv::ColMajor|tmv::RowMajor>   (TMV_Matrix.h)
//   tmv::MatrixView / ConstMatrixView: view(), transpose(), conjugate(),
//   adjoint(), subMatrix(i1,i2,j1,j2[,istep,jstep]), colRange, rowRange
//                                                  (TMV_Matrix.h:1280-1600,
//                                                   TMV_BaseMatrix.h)
//   C = A*B, C = x*A*B, C += A*B, C -= A*B, C += x*A*B, m *= m, m *= x,
//   MultMM<add>(alpha, A, B, C) on matrices and views, aliasing allowed
//                                                  (TMV_MatrixArith.h:700-812,
//                                                   TMV_MatrixArithFunc.h:53-57,
//                                                   src/TMV_MultMM.cpp:430-453)
//   x = b/m, X = B/m (LDiv), x = b%m (RDiv)        (TMV_MatrixArith.h:979-1177)
//   m.divideUsing(tmv::LU | tmv::QR), m.saveDiv(), m.resetDiv(), m.unsetDiv(),
//   m.setDiv(), m.lud() -> LUDiv<T> with getL/getU/getP/det/logDet/
//   isSingular/isTrans                             (TMV_Matrix.h:933-975,
//                                                   TMV_LUD.h:90-161)
//   complex matrices: Matrix<std::complex<double|float>> with *, MultMM,
//   x = b/m, b%m                                   (src/TMV_MultMM_Block.cpp:
//                                                   334-392; LU as above)
//   m.inverse(), m.det(), m.logDet(&sign)          (TMV_Matrix.h:869-871,
//                                                   926-928,
//                                                   src/TMV_LUD.cpp:133-169)
//   LU_Decompose(Matrix& | MatrixView, Permutation&) on column-major,
//   row-major, strided and conjugated views       (TMV_LUD.h:79-88,
//                                                   src/TMV_LUDecompose.cpp:
//                                                   514-538)
//   LUDiv::checkDecomp, LUDiv::isSingular (diagonal scanned on the device)
//                                                  (src/TMV_LUD.cpp:195-254)
// Products stage each operand to the GPU once (a view with unit stride in
// either direction goes as it is, any other view is packed first, as TMV
// copies a view that is not BLAS-compatible) and write C back only after
// the product is complete, so C may alias A or B (m *= m).
// Like TMV, the factorisation is made lazily by the first division, reused
// while saveDiv() is set and discarded after each division otherwise
// (src/TMV_BaseMatrix.cpp:215-274); a row-major matrix is factorised as its
// transpose (isTrans, src/TMV_LUD.cpp:56-69).  Data live in host memory as
// in TMV; each operation stages them to the GPU.  Errors throw tmv::Error /
// tmv::Singular as TMV does; status codes never escape.
//
// Link with libtmvx.so and libamdhip64 (HIP runtime for the staging copies).
#ifndef TMVX_TMV_H
#define TMVX_TMV_H

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "tmvx.h"

namespace tmv {

enum StorageType { ColMajor = 0, RowMajor = 2 };
enum DivType { XX = 0, LU = 1, QR = 4 };    // TMV's values (TMV_Base.h:209-210)

class Error : public std::runtime_error {
public:
    explicit Error(const std::string& s) : std::runtime_error("TMV Error: " + s) {}
};

class Singular : public Error {
public:
    explicit Singular(const std::string& s) : Error("Singular matrix: " + s) {}
};

namespace detail {

inline void check(int rc, const char* where)
{
    if (rc == TMVX_OK) return;
    if (rc == TMVX_ERR_SINGULAR) throw Singular(where);
    throw Error(std::string(where) + ": " + tmvx_status_str(rc));
}

inline void hcheck(hipError_t e, const char* where)
{
    if (e != hipSuccess) throw Error(std::string(where) + ": " + hipGetErrorString(e));
}

// Device buffer (RAII).
template <class T> class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { if (p_) (void)hipFree(p_); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void resize(size_t n)
    {
        if (n == n_) return;
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = n;
        if (n) hcheck(hipMalloc((void**)&p_, n * sizeof(T)), "hipMalloc");
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T* h, size_t n) { hcheck(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "upload"); }
    void download(T* h, size_t n) const { hcheck(hipMemcpy(h, p_, n * sizeof(T), hipMemcpyDeviceToHost), "download"); }
    // `cols` runs of `rows` contiguous elements, `ld` apart on the host,
    // packed densely (rows apart) on the device.
    void upload2d(const T* h, ptrdiff_t ld, ptrdiff_t rows, ptrdiff_t cols)
    {
        hcheck(hipMemcpy2D(p_, rows * sizeof(T), h, ld * sizeof(T), rows * sizeof(T), cols,
                           hipMemcpyHostToDevice), "upload");
    }
    void download2d(T* h, ptrdiff_t ld, ptrdiff_t rows, ptrdiff_t cols) const
    {
        hcheck(hipMemcpy2D(h, ld * sizeof(T), p_, rows * sizeof(T), rows * sizeof(T), cols,
                           hipMemcpyDeviceToHost), "download");
    }
private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

template <class T> struct is_complex : std::false_type {};
template <class R> struct is_complex<std::complex<R>> : std::true_type {};
template <class T> inline T conj_if(const T& x, bool) { return x; }
template <class R> inline std::complex<R> conj_if(const std::complex<R>& x, bool c)
{ return c ? std::conj(x) : x; }
// TMV's Abs2: |x| for real x, |re| + |im| for complex x.
template <class T> inline double abs2(const T& x) { return std::fabs((double)x); }
template <class R> inline double abs2(const std::complex<R>& x)
{ return std::fabs((double)x.real()) + std::fabs((double)x.imag()); }
template <class T> inline double absv(const T& x) { return std::abs(x); }

template <class T> struct Abi;
template <> struct Abi<double> {
    static int gemm(int la, int lb, int, int, int64_t m, int64_t n, int64_t k, double al,
                    const double* A, int64_t lda, const double* B, int64_t ldb, double be,
                    double* C, int64_t ldc)
    { return tmvx_gemm_f64(la, lb, m, n, k, al, A, lda, B, ldb, be, C, ldc, nullptr); }
    static int diag_range(int64_t n, const double* LU, int64_t lda, double* mn, double* mx)
    { return tmvx_lu_diag_range_f64(n, LU, lda, mn, mx, nullptr); }
    static int getrf(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_f64(n, A, lda, piv, info, nullptr); }
    static int getrf_mn(int64_t m, int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_mn_f64(m, n, A, lda, piv, info, nullptr); }
    static int getrs(int tr, int64_t n, int64_t nr, const double* LU, int64_t lda, const int64_t* p,
                     double* B, int64_t ldb)
    { return tmvx_getrs_f64(tr, n, nr, LU, lda, p, B, ldb, 1, nullptr); }
    static int geqrf(int64_t m, int64_t n, double* A, int64_t lda, double* beta)
    { return tmvx_geqrf_f64(m, n, A, lda, beta, nullptr); }
    static int geqrs(int64_t m, int64_t n, int64_t nr, const double* QR, int64_t lda,
                     const double* beta, double* B, int64_t ldb)
    { return tmvx_geqrs_f64(m, n, nr, QR, lda, beta, B, ldb, nullptr); }
    static int getri(int64_t n, const double* LU, int64_t lda, const int64_t* p, double* X,
                     int64_t ldx)
    { return tmvx_getri_f64(n, LU, lda, p, X, ldx, nullptr); }
    static int logdet(int64_t n, const double* LU, int64_t lda, const int64_t* p, double* ld,
                      double* sg)
    { return tmvx_logdet_f64(n, LU, lda, p, ld, sg, nullptr); }
};
template <> struct Abi<float> {
    static int gemm(int la, int lb, int, int, int64_t m, int64_t n, int64_t k, float al,
                    const float* A, int64_t lda, const float* B, int64_t ldb, float be, float* C,
                    int64_t ldc)
    { return tmvx_gemm_f32(la, lb, m, n, k, al, A, lda, B, ldb, be, C, ldc, nullptr); }
    static int diag_range(int64_t n, const float* LU, int64_t lda, double* mn, double* mx)
    { return tmvx_lu_diag_range_f32(n, LU, lda, mn, mx, nullptr); }
    static int getrf(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_f32(n, A, lda, piv, info, nullptr); }
    static int getrf_mn(int64_t m, int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_mn_f32(m, n, A, lda, piv, info, nullptr); }
    static int getrs(int tr, int64_t n, int64_t nr, const float* LU, int64_t lda, const int64_t* p,
                     float* B, int64_t ldb)
    { return tmvx_getrs_f32(tr, n, nr, LU, lda, p, B, ldb, 1, nullptr); }
    static int geqrf(int64_t m, int64_t n, float* A, int64_t lda, float* beta)
    { return tmvx_geqrf_f32(m, n, A, lda, beta, nullptr); }
    static int geqrs(int64_t m, int64_t n, int64_t nr, const float* QR, int64_t lda,
                     const float* beta, float* B, int64_t ldb)
    { return tmvx_geqrs_f32(m, n, nr, QR, lda, beta, B, ldb, nullptr); }
    static int getri(int64_t n, const float* LU, int64_t lda, const int64_t* p, float* X,
                     int64_t ldx)
    { return tmvx_getri_f32(n, LU, lda, p, X, ldx, nullptr); }
    static int logdet(int64_t n, const float* LU, int64_t lda, const int64_t* p, double* ld,
                      double* sg)
    { return tmvx_logdet_f32(n, LU, lda, p, ld, sg, nullptr); }
};

// Complex element types: products and LU division through the complex
// entry points (factors and solves; det/logDet/inverse stay real-only).
template <> struct Abi<std::complex<double>> {
    typedef std::complex<double> Z;
    static int gemm(int la, int lb, int ca, int cb, int64_t m, int64_t n, int64_t k, Z al,
                    const Z* A, int64_t lda, const Z* B, int64_t ldb, Z be, Z* C, int64_t ldc)
    {
        return tmvx_gemm_z(la, lb, ca, cb, m, n, k, al.real(), al.imag(),
                           reinterpret_cast<const double*>(A), lda,
                           reinterpret_cast<const double*>(B), ldb, be.real(), be.imag(),
                           reinterpret_cast<double*>(C), ldc, nullptr);
    }
    static int getrf(int64_t n, Z* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_z(n, reinterpret_cast<double*>(A), lda, piv, info, nullptr); }
    static int getrf_mn(int64_t, int64_t, Z*, int64_t, int64_t*, int64_t*)
    { return TMVX_ERR_UNSUPPORTED; }
    static int diag_range(int64_t n, const Z* LU, int64_t lda, double* mn, double* mx)
    { return tmvx_lu_diag_range_z(n, reinterpret_cast<const double*>(LU), lda, mn, mx, nullptr); }
    static int getri(int64_t, const Z*, int64_t, const int64_t*, Z*, int64_t)
    { return TMVX_ERR_UNSUPPORTED; }
    static int geqrf(int64_t, int64_t, Z*, int64_t, Z*) { return TMVX_ERR_UNSUPPORTED; }
    static int geqrs(int64_t, int64_t, int64_t, const Z*, int64_t, const Z*, Z*, int64_t)
    { return TMVX_ERR_UNSUPPORTED; }
    static int getrs(int tr, int64_t n, int64_t nr, const Z* LU, int64_t lda, const int64_t* p,
                     Z* B, int64_t ldb)
    {
        return tmvx_getrs_z(tr, n, nr, reinterpret_cast<const double*>(LU), lda, p,
                            reinterpret_cast<double*>(B), ldb, nullptr);
    }
};
template <> struct Abi<std::complex<float>> {
    typedef std::complex<float> Z;
    static int gemm(int la, int lb, int ca, int cb, int64_t m, int64_t n, int64_t k, Z al,
                    const Z* A, int64_t lda, const Z* B, int64_t ldb, Z be, Z* C, int64_t ldc)
    {
        return tmvx_gemm_c(la, lb, ca, cb, m, n, k, al.real(), al.imag(),
                           reinterpret_cast<const float*>(A), lda,
                           reinterpret_cast<const float*>(B), ldb, be.real(), be.imag(),
                           reinterpret_cast<float*>(C), ldc, nullptr);
    }
    static int getrf(int64_t n, Z* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_c(n, reinterpret_cast<float*>(A), lda, piv, info, nullptr); }
    static int getrf_mn(int64_t, int64_t, Z*, int64_t, int64_t*, int64_t*)
    { return TMVX_ERR_UNSUPPORTED; }
    static int diag_range(int64_t n, const Z* LU, int64_t lda, double* mn, double* mx)
    { return tmvx_lu_diag_range_c(n, reinterpret_cast<const float*>(LU), lda, mn, mx, nullptr); }
    static int getri(int64_t, const Z*, int64_t, const int64_t*, Z*, int64_t)
    { return TMVX_ERR_UNSUPPORTED; }
    static int geqrf(int64_t, int64_t, Z*, int64_t, Z*) { return TMVX_ERR_UNSUPPORTED; }
    static int geqrs(int64_t, int64_t, int64_t, const Z*, int64_t, const Z*, Z*, int64_t)
    { return TMVX_ERR_UNSUPPORTED; }
    static int getrs(int tr, int64_t n, int64_t nr, const Z* LU, int64_t lda, const int64_t* p,
                     Z* B, int64_t ldb)
    {
        return tmvx_getrs_c(tr, n, nr, reinterpret_cast<const float*>(LU), lda, p,
                            reinterpret_cast<float*>(B), ldb, nullptr);
    }
};

} // namespace detail

// Interchange list as TMV's Permutation stores it (include/tmv/
// TMV_Permutation.h:175-280): P[i] is the row swapped with row i at step i.
class Permutation {
public:
    explicit Permutation(ptrdiff_t n = 0) : p_((size_t)n) { for (ptrdiff_t i = 0; i < n; ++i) p_[(size_t)i] = i; }
    ptrdiff_t size() const { return (ptrdiff_t)p_.size(); }
    const ptrdiff_t* getValues() const { return p_.data(); }
    ptrdiff_t* getMem() { return p_.data(); }
    ptrdiff_t operator[](ptrdiff_t i) const { return p_[(size_t)i]; }
    // (-1)^(number of actual interchanges) (TMV_Permutation.h:273-278)
    int det() const
    {
        int d = 1;
        for (size_t i = 0; i < p_.size(); ++i) if (p_[i] != (ptrdiff_t)i) d = -d;
        return d;
    }
    void resize(ptrdiff_t n) { p_.assign((size_t)n, 0); }
private:
    std::vector<ptrdiff_t> p_;
};

template <class T> class Vector {
public:
    explicit Vector(ptrdiff_t n = 0, T x = T(0)) : v_((size_t)n, x) {}
    Vector(std::initializer_list<T> l) : v_(l) {}
    ptrdiff_t size() const { return (ptrdiff_t)v_.size(); }
    T& operator()(ptrdiff_t i) { return v_[(size_t)i]; }
    const T& operator()(ptrdiff_t i) const { return v_[(size_t)i]; }
    T* ptr() { return v_.data(); }
    const T* cptr() const { return v_.data(); }
private:
    std::vector<T> v_;
};

template <class T> class LUDiv;
template <class T> class QRDiv;
template <class T> class MatrixView;
template <class T> struct ProdMM;
template <class T> struct ProdXM;

// Element of a possibly conjugated view (TMV's VarConjRef): reads and
// writes go through conj() when the view is conjugated.
template <class T> class ConjRef {
public:
    ConjRef(T* p, bool c) : p_(p), c_(c) {}
    operator T() const { return detail::conj_if(*p_, c_); }
    ConjRef& operator=(const T& x) { *p_ = detail::conj_if(x, c_); return *this; }
    ConjRef& operator=(const ConjRef& o) { return *this = T(o); }
    ConjRef& operator+=(const T& x) { return *this = T(*this) + x; }
    ConjRef& operator-=(const T& x) { return *this = T(*this) - x; }
    ConjRef& operator*=(const T& x) { return *this = T(*this) * x; }
private:
    T* p_;
    bool c_;
};

// Read-only view of m x n elements at p + i*stepi + j*stepj, optionally
// conjugated (TMV's ConstMatrixView, include/tmv/TMV_Matrix.h).  Steps may
// be any non-zero values, negative ones included.
template <class T> class ConstMatrixView {
public:
    ConstMatrixView(const T* p, ptrdiff_t m, ptrdiff_t n, ptrdiff_t si, ptrdiff_t sj,
                    bool c = false)
        : p_(p), m_(m), n_(n), si_(si), sj_(sj), c_(c && detail::is_complex<T>::value) {}
    ptrdiff_t colsize() const { return m_; }
    ptrdiff_t rowsize() const { return n_; }
    ptrdiff_t stepi() const { return si_; }
    ptrdiff_t stepj() const { return sj_; }
    bool isconj() const { return c_; }
    bool isSquare() const { return m_ == n_; }
    // unit step down the columns (across the rows) with the other step
    // leaving room for them: storage a GEMM can take as it is
    bool iscm() const { return si_ == 1 && (n_ <= 1 || sj_ >= (m_ > 1 ? m_ : 1)); }
    bool isrm() const { return sj_ == 1 && (m_ <= 1 || si_ >= (n_ > 1 ? n_ : 1)); }
    const T* cptr() const { return p_; }
    T operator()(ptrdiff_t i, ptrdiff_t j) const { return detail::conj_if(p_[i * si_ + j * sj_], c_); }
    ConstMatrixView transpose() const { return ConstMatrixView(p_, n_, m_, sj_, si_, c_); }
    ConstMatrixView conjugate() const { return ConstMatrixView(p_, m_, n_, si_, sj_, !c_); }
    ConstMatrixView adjoint() const { return transpose().conjugate(); }
    ConstMatrixView subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2) const
    { check_range(i1, i2, j1, j2, 1, 1); return ConstMatrixView(p_ + i1 * si_ + j1 * sj_, i2 - i1, j2 - j1, si_, sj_, c_); }
    // every istep-th row and jstep-th column of [i1, i2) x [j1, j2)
    ConstMatrixView subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2,
                              ptrdiff_t istep, ptrdiff_t jstep) const
    {
        check_range(i1, i2, j1, j2, istep, jstep);
        return ConstMatrixView(p_ + i1 * si_ + j1 * sj_, (i2 - i1) / istep, (j2 - j1) / jstep,
                               si_ * istep, sj_ * jstep, c_);
    }
    ConstMatrixView colRange(ptrdiff_t j1, ptrdiff_t j2) const { return subMatrix(0, m_, j1, j2); }
    ConstMatrixView rowRange(ptrdiff_t i1, ptrdiff_t i2) const { return subMatrix(i1, i2, 0, n_); }
protected:
    void check_range(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2, ptrdiff_t is,
                     ptrdiff_t js) const
    {
        if (is <= 0 || js <= 0 || i1 < 0 || j1 < 0 || i2 > m_ || j2 > n_ || i1 > i2 || j1 > j2 ||
            (i2 - i1) % is || (j2 - j1) % js)
            throw Error("subMatrix: range outside the matrix");
    }
    const T* p_;
    ptrdiff_t m_, n_, si_, sj_;
    bool c_;
};

// Writable view (TMV's MatrixView).  Assignment copies the elements of
// the source (through a temporary, so overlapping views are safe); the
// arithmetic assignments run on the GPU.
template <class T> class MatrixView : public ConstMatrixView<T> {
    typedef ConstMatrixView<T> Base;
public:
    MatrixView(T* p, ptrdiff_t m, ptrdiff_t n, ptrdiff_t si, ptrdiff_t sj, bool c = false)
        : Base(p, m, n, si, sj, c) {}
    T* ptr() const { return const_cast<T*>(this->p_); }
    ConjRef<T> operator()(ptrdiff_t i, ptrdiff_t j) const
    { return ConjRef<T>(ptr() + i * this->si_ + j * this->sj_, this->c_); }
    MatrixView transpose() const { return MatrixView(ptr(), this->n_, this->m_, this->sj_, this->si_, this->c_); }
    MatrixView conjugate() const { return MatrixView(ptr(), this->m_, this->n_, this->si_, this->sj_, !this->c_); }
    MatrixView adjoint() const { return transpose().conjugate(); }
    MatrixView subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2) const
    { return subMatrix(i1, i2, j1, j2, 1, 1); }
    MatrixView subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2, ptrdiff_t istep,
                         ptrdiff_t jstep) const
    {
        this->check_range(i1, i2, j1, j2, istep, jstep);
        return MatrixView(ptr() + i1 * this->si_ + j1 * this->sj_, (i2 - i1) / istep,
                          (j2 - j1) / jstep, this->si_ * istep, this->sj_ * jstep, this->c_);
    }
    MatrixView colRange(ptrdiff_t j1, ptrdiff_t j2) const { return subMatrix(0, this->m_, j1, j2); }
    MatrixView rowRange(ptrdiff_t i1, ptrdiff_t i2) const { return subMatrix(i1, i2, 0, this->n_); }

    const MatrixView& operator=(const ConstMatrixView<T>& o) const
    {
        if (o.colsize() != this->m_ || o.rowsize() != this->n_) throw Error("view =: size mismatch");
        std::vector<T> tmp((size_t)(this->m_ * this->n_));
        for (ptrdiff_t j = 0; j < this->n_; ++j)
            for (ptrdiff_t i = 0; i < this->m_; ++i) tmp[(size_t)(i + j * this->m_)] = o(i, j);
        for (ptrdiff_t j = 0; j < this->n_; ++j)
            for (ptrdiff_t i = 0; i < this->m_; ++i) (*this)(i, j) = tmp[(size_t)(i + j * this->m_)];
        return *this;
    }
    const MatrixView& operator=(const MatrixView& o) const { return *this = static_cast<const Base&>(o); }
    const MatrixView& setZero() const
    {
        for (ptrdiff_t j = 0; j < this->n_; ++j)
            for (ptrdiff_t i = 0; i < this->m_; ++i) (*this)(i, j) = T(0);
        return *this;
    }
    const MatrixView& operator*=(T x) const
    {
        for (ptrdiff_t j = 0; j < this->n_; ++j)
            for (ptrdiff_t i = 0; i < this->m_; ++i) (*this)(i, j) *= x;
        return *this;
    }
    inline const MatrixView& operator=(const ProdMM<T>& p) const;
    inline const MatrixView& operator+=(const ProdMM<T>& p) const;
    inline const MatrixView& operator-=(const ProdMM<T>& p) const;
    inline const MatrixView& operator*=(const ConstMatrixView<T>& m) const;
};

template <class T, int S = ColMajor> class Matrix {
public:
    Matrix() : m_(0), n_(0) {}
    Matrix(ptrdiff_t m, ptrdiff_t n, T x = T(0)) : m_(m), n_(n), a_((size_t)(m * n), x) {}
    // Row-major element list, as TMV's `m << a, b, c` (TMV_ListInit.h).
    Matrix(ptrdiff_t m, ptrdiff_t n, std::initializer_list<T> rowlist) : Matrix(m, n)
    {
        if ((ptrdiff_t)rowlist.size() != m * n) throw Error("list size");
        ptrdiff_t k = 0;
        for (T x : rowlist) { (*this)(k / n, k % n) = x; ++k; }
    }
    Matrix(const Matrix& o) : m_(o.m_), n_(o.n_), a_(o.a_), divtype_(o.divtype_) {}
    template <int S2> explicit Matrix(const Matrix<T, S2>& o) : Matrix(o.colsize(), o.rowsize())
    { for (ptrdiff_t j = 0; j < n_; ++j) for (ptrdiff_t i = 0; i < m_; ++i) (*this)(i, j) = o(i, j); }
    Matrix& operator=(const Matrix& o)
    {
        if (this != &o) { m_ = o.m_; n_ = o.n_; a_ = o.a_; resetDiv(); }
        return *this;
    }

    ptrdiff_t colsize() const { return m_; }
    ptrdiff_t rowsize() const { return n_; }
    ptrdiff_t nrows() const { return m_; }
    ptrdiff_t ncols() const { return n_; }
    bool isSquare() const { return m_ == n_; }
    static constexpr bool isrm() { return S == RowMajor; }
    static constexpr bool iscm() { return S == ColMajor; }
    ptrdiff_t stepi() const { return isrm() ? n_ : 1; }
    ptrdiff_t stepj() const { return isrm() ? 1 : m_; }
    T& operator()(ptrdiff_t i, ptrdiff_t j) { return a_[(size_t)(i * stepi() + j * stepj())]; }
    const T& operator()(ptrdiff_t i, ptrdiff_t j) const { return a_[(size_t)(i * stepi() + j * stepj())]; }
    T* ptr() { return a_.data(); }
    const T* cptr() const { return a_.data(); }
    void setZero() { std::fill(a_.begin(), a_.end(), T(0)); resetDiv(); }

    // ---- views (TMV_Matrix.h: view, transpose, conjugate, adjoint,
    //      subMatrix, colRange, rowRange) ----------------------------------
    MatrixView<T> view() { return MatrixView<T>(ptr(), m_, n_, stepi(), stepj()); }
    ConstMatrixView<T> view() const { return ConstMatrixView<T>(cptr(), m_, n_, stepi(), stepj()); }
    operator ConstMatrixView<T>() const { return view(); }
    MatrixView<T> transpose() { return view().transpose(); }
    ConstMatrixView<T> transpose() const { return view().transpose(); }
    MatrixView<T> conjugate() { return view().conjugate(); }
    ConstMatrixView<T> conjugate() const { return view().conjugate(); }
    MatrixView<T> adjoint() { return view().adjoint(); }
    ConstMatrixView<T> adjoint() const { return view().adjoint(); }
    MatrixView<T> subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2)
    { return view().subMatrix(i1, i2, j1, j2); }
    ConstMatrixView<T> subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2) const
    { return view().subMatrix(i1, i2, j1, j2); }
    MatrixView<T> subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2,
                            ptrdiff_t is, ptrdiff_t js)
    { return view().subMatrix(i1, i2, j1, j2, is, js); }
    ConstMatrixView<T> subMatrix(ptrdiff_t i1, ptrdiff_t i2, ptrdiff_t j1, ptrdiff_t j2,
                                 ptrdiff_t is, ptrdiff_t js) const
    { return view().subMatrix(i1, i2, j1, j2, is, js); }
    MatrixView<T> colRange(ptrdiff_t j1, ptrdiff_t j2) { return view().colRange(j1, j2); }
    ConstMatrixView<T> colRange(ptrdiff_t j1, ptrdiff_t j2) const { return view().colRange(j1, j2); }
    MatrixView<T> rowRange(ptrdiff_t i1, ptrdiff_t i2) { return view().rowRange(i1, i2); }
    ConstMatrixView<T> rowRange(ptrdiff_t i1, ptrdiff_t i2) const { return view().rowRange(i1, i2); }

    // a copy of the view's elements
    Matrix(const ConstMatrixView<T>& v) : Matrix(v.colsize(), v.rowsize())
    { for (ptrdiff_t j = 0; j < n_; ++j) for (ptrdiff_t i = 0; i < m_; ++i) (*this)(i, j) = v(i, j); }
    Matrix(const MatrixView<T>& v) : Matrix(static_cast<const ConstMatrixView<T>&>(v)) {}
    Matrix& operator=(const ConstMatrixView<T>& v)
    {
        Matrix t(v);                      // v may look into this matrix
        a_.swap(t.a_); m_ = t.m_; n_ = t.n_;
        resetDiv();
        return *this;
    }
    // products: C = A*B and friends, evaluated on the GPU
    inline Matrix(const ProdMM<T>& p);
    inline Matrix(const ProdXM<T>& p);
    inline Matrix& operator=(const ProdMM<T>& p);
    inline Matrix& operator+=(const ProdMM<T>& p);
    inline Matrix& operator-=(const ProdMM<T>& p);
    inline Matrix& operator*=(const ConstMatrixView<T>& m);     // m = m * x (aliasing allowed)
    Matrix& operator*=(T x)
    {
        for (T& v : a_) v *= x;
        resetDiv();
        return *this;
    }

    // ---- division control (TMV_Matrix.h:933-975, TMV_BaseMatrix.cpp) ----
    void divideUsing(DivType dt)
    {
        if (dt != LU && dt != QR && dt != XX)
            throw Error("only LU and QR division are provided by this library");
        if (dt != divtype_) unsetDiv();
        divtype_ = dt;
    }
    // XX resolves to LU for a square matrix, QR otherwise
    // (DivHelper::resetDivType, src/TMV_BaseMatrix.cpp:234-236).
    bool usesQR() const { return divtype_ == QR || (divtype_ == XX && !isSquare()); }
    void saveDiv() { savediv_ = true; }
    void setDiv() const
    {
        if (usesQR()) { if (!qdiv_) qdiv_.reset(new QRDiv<T>(*this)); }
        else if (!div_) div_.reset(new LUDiv<T>(*this));
    }
    void unsetDiv() const { div_.reset(); qdiv_.reset(); }
    void resetDiv() const { unsetDiv(); }
    bool divIsSet() const { return (bool)div_ || (bool)qdiv_; }
    QRDiv<T>& qrd() const
    {
        if (!usesQR()) throw Error("qrd(): the division type is not QR");
        setDiv();
        return *qdiv_;
    }
    LUDiv<T>& lud() const
    {
        if (usesQR()) throw Error("lud(): the division type is QR");
        setDiv();
        return *div_;
    }
    // Divisions release the factorisation unless saveDiv() was called
    // (src/TMV_BaseMatrix.cpp:273-274).
    void doneDiv() const { if (!savediv_) unsetDiv(); }

    T det() const
    {
        if (m_ == 0) return T(1);
        if (usesQR()) throw Error("det() is provided with LU division");
        setDiv();
        const T d = div_->det();
        doneDiv();
        return d;
    }
    T logDet(T* sign = nullptr) const
    {
        if (m_ == 0) { if (sign) *sign = T(1); return T(0); }
        if (usesQR()) throw Error("logDet() is provided with LU division");
        setDiv();
        const T ld = div_->logDet(sign);
        doneDiv();
        return ld;
    }
    // m.inverse() with LU division (TMV_Matrix.h:926-928, LUDiv::
    // makeInverse): a new matrix of the same storage order.
    Matrix inverse() const
    {
        if (!isSquare()) throw Error("inverse of a non-square matrix needs QR or SV division");
        if (usesQR()) throw Error("inverse() is provided with LU division");
        Matrix R(m_, n_);
        if (m_ == 0) return R;
        setDiv();
        div_->makeInverse(R.ptr());
        doneDiv();
        return R;
    }

private:
    ptrdiff_t m_, n_;
    std::vector<T> a_;
    DivType divtype_ = XX;
    bool savediv_ = false;
    mutable std::shared_ptr<LUDiv<T>> div_;
    mutable std::shared_ptr<QRDiv<T>> qdiv_;
};

// ---------------------------------------------------------------------------
// QR divider (QRDiv, src/TMV_QRD.cpp; QR_LDiv, src/TMV_QRDiv.cpp:33-94):
// A (m x n, m >= n) factored on the device; x = b/A is the least-squares
// solution for m > n.
// ---------------------------------------------------------------------------
template <class T> class QRDiv {
public:
    template <int S>
    explicit QRDiv(const Matrix<T, S>& A) : m_(A.colsize()), n_(A.rowsize())
    {
        if (m_ < n_) throw Error("QR division needs colsize >= rowsize");
        qr_.resize((size_t)(m_ * n_));
        beta_.resize((size_t)n_);
        if (n_ == 0) return;
        const Matrix<T, ColMajor> Ac(A);
        qr_.upload(Ac.cptr(), (size_t)(m_ * n_));
        detail::check(detail::Abi<T>::geqrf(m_, n_, qr_.get(), m_, beta_.get()), "QR_Decompose");
    }
    // x (n x nr, column-major) from b (m x nr, column-major).
    void lDivCM(const T* b, T* x, ptrdiff_t nr) const
    {
        if (n_ == 0 || nr == 0) return;
        detail::DevBuf<T> bb((size_t)(m_ * nr));
        bb.upload(b, (size_t)(m_ * nr));
        detail::check(detail::Abi<T>::geqrs(m_, n_, nr, qr_.get(), m_, beta_.get(), bb.get(), m_),
                      "QR_LDiv");
        std::vector<T> h((size_t)(m_ * nr));
        bb.download(h.data(), (size_t)(m_ * nr));
        for (ptrdiff_t j = 0; j < nr; ++j)
            for (ptrdiff_t i = 0; i < n_; ++i) x[i + j * n_] = h[(size_t)(i + j * m_)];
    }
    // Packed factors: R on and above the diagonal, reflectors below.
    Matrix<T, ColMajor> getQR() const
    {
        Matrix<T, ColMajor> M(m_, n_);
        if (m_ * n_) qr_.download(M.ptr(), (size_t)(m_ * n_));
        return M;
    }
    std::vector<T> getBeta() const
    {
        std::vector<T> b((size_t)n_);
        if (n_) beta_.download(b.data(), (size_t)n_);
        return b;
    }

private:
    ptrdiff_t m_, n_;
    detail::DevBuf<T> qr_;
    detail::DevBuf<T> beta_;
};

// ---------------------------------------------------------------------------
// LU divider: factors live on the device (and, on request, on the host).
// ---------------------------------------------------------------------------
template <class T> class LUDiv {
public:
    template <int S>
    explicit LUDiv(const Matrix<T, S>& A) : LUDiv(A.view()) {}
    // TMV factorises a row-major matrix as its transpose (istrans,
    // src/TMV_LUD.cpp:56-69) and copies any other view first.
    explicit LUDiv(const ConstMatrixView<T>& A)
        : n_(A.colsize()), istrans_(A.isrm() && !A.iscm() && !A.isconj()), P_(A.colsize())
    {
        if (!A.isSquare()) throw Error("LU division requires a square matrix");
        lu_.resize((size_t)(n_ * n_));
        piv_.resize((size_t)n_);
        if (n_ == 0) return;
        if (A.iscm() && !A.isconj()) {
            lu_.upload2d(A.cptr(), n_ > 1 ? A.stepj() : n_, n_, n_);
        } else if (istrans_) {
            // the storage read column-major is A^T
            lu_.upload2d(A.cptr(), n_ > 1 ? A.stepi() : n_, n_, n_);
        } else {
            const Matrix<T, ColMajor> Ac(A);
            lu_.upload(Ac.cptr(), (size_t)(n_ * n_));
        }
        int64_t info = 0;
        detail::check(detail::Abi<T>::getrf(n_, lu_.get(), n_, piv_.get(), &info), "LU_Decompose");
        std::vector<int64_t> hp((size_t)n_);
        piv_.download(hp.data(), (size_t)n_);
        for (ptrdiff_t i = 0; i < n_; ++i) P_.getMem()[i] = (ptrdiff_t)hp[(size_t)i];
    }

    bool isTrans() const { return istrans_; }
    const Permutation& getP() const { return P_; }
    // Packed factors (column-major n x n): L unit-lower, U upper.
    Matrix<T, ColMajor> getLU() const
    {
        Matrix<T, ColMajor> M(n_, n_);
        if (n_) lu_.download(M.ptr(), (size_t)(n_ * n_));
        return M;
    }
    Matrix<T, ColMajor> getL() const
    {
        Matrix<T, ColMajor> M = getLU();
        for (ptrdiff_t j = 0; j < n_; ++j)
            for (ptrdiff_t i = 0; i <= j; ++i) M(i, j) = (i == j) ? T(1) : T(0);
        return M;
    }
    Matrix<T, ColMajor> getU() const
    {
        Matrix<T, ColMajor> M = getLU();
        for (ptrdiff_t j = 0; j < n_; ++j)
            for (ptrdiff_t i = j + 1; i < n_; ++i) M(i, j) = T(0);
        return M;
    }
    // log|det| and sign (src/TMV_LUD.cpp:145-156), reduced on the device;
    // the factors of A^T have the same determinant.
    T logDet(T* sign = nullptr) const
    {
        double ld = 0.0, sg = 1.0;
        if (n_ > 0)
            detail::check(detail::Abi<T>::logdet(n_, lu_.get(), n_, piv_.get(), &ld, &sg),
                          "LUDiv::logDet");
        if (sign) *sign = T(sg);
        return T(ld);
    }
    // A^-1 written into n x n storage laid out like A (src/TMV_LUD.cpp:
    // 158-169).  For a row-major A the factors are those of A^T, whose
    // column-major inverse is A^-1 in row-major order, as TMV's
    // LU_Inverse(LUx, P, minv.transpose()) forms it.
    void makeInverse(T* out) const
    {
        if (n_ == 0) return;
        detail::DevBuf<T> x((size_t)(n_ * n_));
        detail::check(detail::Abi<T>::getri(n_, lu_.get(), n_, piv_.get(), x.get(), n_),
                      "LU_Inverse");
        x.download(out, (size_t)(n_ * n_));
    }
    T det() const
    {
        T sg;
        const T ld = logDet(&sg);
        return sg == T(0) ? T(0) : sg * std::exp(ld);
    }
    // min Abs2(U_ii) <= eps * max Abs2(U_ii) (src/TMV_LUD.cpp:195-200),
    // the diagonal scanned on the device
    bool isSingular() const
    {
        if (n_ == 0) return false;
        double mn = 0.0, mx = 0.0;
        detail::check(detail::Abi<T>::diag_range(n_, lu_.get(), n_, &mn, &mx), "LUDiv::isSingular");
        typedef decltype(std::abs(T(0))) RT;
        return mn <= (double)std::numeric_limits<RT>::epsilon() * mx;
    }

    // TMV's LUDiv::checkDecomp (src/TMV_LUD.cpp:225-254): P L U against
    // the matrix (the product L U on the GPU), ||PLU - M||_F /
    // (||L||_F ||U||_F) below cond(M) n eps.  cond(M) here is the 1-norm
    // condition ||M||_1 ||M^-1||_1 from the device inverse (TMV takes the
    // 2-norm condition of an SVD; the two agree within a factor n).
    template <int S>
    bool checkDecomp(const Matrix<T, S>& m, std::ostream* fout = nullptr) const
    {
        typedef decltype(std::abs(T(0))) RT;
        if (m.colsize() != n_ || m.rowsize() != n_) throw Error("checkDecomp: size mismatch");
        if (n_ == 0) return true;
        const Matrix<T, ColMajor> L = getL(), U = getU();
        Matrix<T, ColMajor> plu = L * U;
        // rows of L U in the interchanged order: undo the swaps, last first
        for (ptrdiff_t i = n_ - 1; i >= 0; --i) {
            const ptrdiff_t p = P_[i];
            if (p != i) for (ptrdiff_t j = 0; j < n_; ++j) std::swap(plu(i, j), plu(p, j));
        }
        // the factors are those of M^T when isTrans
        auto mel = [&](ptrdiff_t i, ptrdiff_t j) { return istrans_ ? m(j, i) : m(i, j); };
        double d2 = 0, l2 = 0, u2 = 0;
        for (ptrdiff_t j = 0; j < n_; ++j)
            for (ptrdiff_t i = 0; i < n_; ++i) {
                const double e = detail::absv(plu(i, j) - mel(i, j));
                d2 += e * e;
                l2 += std::norm(L(i, j));
                u2 += std::norm(U(i, j));
            }
        const double nm = std::sqrt(d2) / (std::sqrt(l2) * std::sqrt(u2));
        const double cond = condition1(m);
        if (fout) {
            *fout << "LUDiv: isTrans = " << istrans_ << ", interchanges:";
            for (ptrdiff_t i = 0; i < n_; ++i) *fout << ' ' << P_[i];
            *fout << "\nNorm(M-PLU)/(Norm(L)Norm(U)) = " << nm << ", cond = " << cond << '\n';
        }
        return nm < cond * (double)n_ * (double)std::numeric_limits<RT>::epsilon();
    }

    // m := A^-1 m (LDivEq), nr right-hand sides stored column-major.
    void lDivEqCM(T* m, ptrdiff_t nr) const { solve(m, nr, false); }
    // m := m A^-1 for row vectors, i.e. A^T y = m (RDivEq).
    void rDivEqCM(T* m, ptrdiff_t nr) const { solve(m, nr, true); }

private:
    template <int S>
    double condition1(const Matrix<T, S>& m) const
    {
        double na = 0;
        for (ptrdiff_t j = 0; j < n_; ++j) {
            double c = 0;
            for (ptrdiff_t i = 0; i < n_; ++i) c += detail::absv(m(i, j));
            na = std::fmax(na, c);
        }
        std::vector<T> inv((size_t)(n_ * n_));
        if (detail::is_complex<T>::value) {
            // columns of M^-1 from the solve (no complex inverse entry)
            for (ptrdiff_t j = 0; j < n_; ++j) inv[(size_t)(j + j * n_)] = T(1);
            solve(inv.data(), n_, false);
        } else {
            makeInverse(inv.data());        // laid out like the factors' matrix
        }
        double ni = 0;
        for (ptrdiff_t j = 0; j < n_; ++j) {
            double c = 0;
            for (ptrdiff_t i = 0; i < n_; ++i)
                c += detail::absv(istrans_ && !detail::is_complex<T>::value
                                      ? inv[(size_t)(j + i * n_)] : inv[(size_t)(i + j * n_)]);
            ni = std::fmax(ni, c);
        }
        return na * ni;
    }

    void solve(T* m, ptrdiff_t nr, bool right) const
    {
        if (n_ == 0 || nr == 0) return;
        detail::DevBuf<T> b((size_t)(n_ * nr));
        b.upload(m, (size_t)(n_ * nr));
        // LUx holds the factors of A (or of A^T when istrans): the left
        // division uses the transposed solve for istrans (src/TMV_LUD.cpp:
        // 92-110), the right division the other way round.
        const int trans = (right != istrans_) ? 1 : 0;
        detail::check(detail::Abi<T>::getrs(trans, n_, nr, lu_.get(), n_, piv_.get(), b.get(), n_),
                      "LU_LDivEq");
        b.download(m, (size_t)(n_ * nr));
    }

    ptrdiff_t n_;
    bool istrans_;
    Permutation P_;
    detail::DevBuf<T> lu_;
    detail::DevBuf<int64_t> piv_;
};

// ---------------------------------------------------------------------------
// GEMM: MultMM<add>(alpha, A, B, C): C (+)= alpha A B (MatrixArithFunc.h:53)
// ---------------------------------------------------------------------------
namespace detail {

// An operand on the device: dense, column- or row-major as the view was
// (a view with unit stride in neither direction is packed column-major
// first, as TMV copies a view that is not BLAS-compatible,
// src/TMV_MultMM.cpp:430-453); the conjugation stays a flag for the GEMM.
template <class T> struct Staged {
    DevBuf<T> buf;
    int layout = TMVX_COLMAJOR;
    int64_t ld = 1;
    int conj = 0;
    explicit Staged(const ConstMatrixView<T>& v)
    {
        const ptrdiff_t m = v.colsize(), n = v.rowsize();
        buf.resize((size_t)(m * n));
        if (m == 0 || n == 0) return;
        if (v.iscm()) {
            buf.upload2d(v.cptr(), n > 1 ? v.stepj() : m, m, n);
            ld = m;
            conj = v.isconj();
        } else if (v.isrm()) {
            buf.upload2d(v.cptr(), m > 1 ? v.stepi() : n, n, m);
            layout = TMVX_ROWMAJOR;
            ld = n;
            conj = v.isconj();
        } else {
            std::vector<T> h((size_t)(m * n));
            for (ptrdiff_t j = 0; j < n; ++j)
                for (ptrdiff_t i = 0; i < m; ++i) h[(size_t)(i + j * m)] = v(i, j);
            buf.upload(h.data(), h.size());
            ld = m;
        }
    }
};

template <bool add, class T>
void multmm(T alpha, const ConstMatrixView<T>& A, const ConstMatrixView<T>& B,
            const MatrixView<T>& C)
{
    if (A.colsize() != C.colsize() || A.rowsize() != B.colsize() || B.rowsize() != C.rowsize())
        throw Error("MultMM: size mismatch");
    const ptrdiff_t m = C.colsize(), n = C.rowsize(), k = A.rowsize();
    if (m == 0 || n == 0) return;
    // Both operands go to the device before C is touched: C may alias them.
    const Staged<T> a(A), b(B);
    DevBuf<T> c((size_t)(m * n));
    const T beta = add ? T(1) : T(0);
    if (!C.isconj() && C.iscm()) {
        const ptrdiff_t ldh = n > 1 ? C.stepj() : m;
        if (add) c.upload2d(C.cptr(), ldh, m, n);
        check(Abi<T>::gemm(a.layout, b.layout, a.conj, b.conj, m, n, k, alpha, a.buf.get(), a.ld,
                           b.buf.get(), b.ld, beta, c.get(), m), "MultMM");
        c.download2d(C.ptr(), ldh, m, n);
    } else if (!C.isconj() && C.isrm()) {
        // C^T (n x m, column-major) = B^T A^T: the operand layouts flip
        // (src/TMV_MultMM.cpp:447-448).
        const ptrdiff_t ldh = m > 1 ? C.stepi() : n;
        if (add) c.upload2d(C.cptr(), ldh, n, m);
        check(Abi<T>::gemm(b.layout ^ 1, a.layout ^ 1, b.conj, a.conj, n, m, k, alpha,
                           b.buf.get(), b.ld, a.buf.get(), a.ld, beta, c.get(), n), "MultMM");
        c.download2d(C.ptr(), ldh, n, m);
    } else {
        // strided or conjugated C: through a packed column-major copy
        std::vector<T> h((size_t)(m * n), T(0));
        if (add)
            for (ptrdiff_t j = 0; j < n; ++j)
                for (ptrdiff_t i = 0; i < m; ++i) h[(size_t)(i + j * m)] = C(i, j);
        if (add) c.upload(h.data(), h.size());
        check(Abi<T>::gemm(a.layout, b.layout, a.conj, b.conj, m, n, k, alpha, a.buf.get(), a.ld,
                           b.buf.get(), b.ld, beta, c.get(), m), "MultMM");
        c.download(h.data(), h.size());
        for (ptrdiff_t j = 0; j < n; ++j)
            for (ptrdiff_t i = 0; i < m; ++i) C(i, j) = h[(size_t)(i + j * m)];
    }
}

template <class T> inline ConstMatrixView<T> cview(const ConstMatrixView<T>& v) { return v; }
template <class T, int S> inline ConstMatrixView<T> cview(const Matrix<T, S>& m) { return m.view(); }
template <class T> inline MatrixView<T> mview(const MatrixView<T>& v) { return v; }
template <class T, int S> inline MatrixView<T> mview(Matrix<T, S>& m) { return m.view(); }

} // namespace detail

template <bool add, class T, int SA, int SB, int SC>
void MultMM(T alpha, const Matrix<T, SA>& A, const Matrix<T, SB>& B, Matrix<T, SC>& C)
{ detail::multmm<add, T>(alpha, A.view(), B.view(), C.view()); }

template <bool add, class T>
void MultMM(T alpha, const ConstMatrixView<T>& A, const ConstMatrixView<T>& B,
            const MatrixView<T>& C)
{ detail::multmm<add, T>(alpha, A, B, C); }

// x*A (TMV's ProdXM) and x*A*B (ProdMM): evaluated when assigned.  A
// product of products evaluates the inner one into a matrix it keeps.
template <class T> struct ProdXM {
    T x;
    ConstMatrixView<T> m;
    std::shared_ptr<const Matrix<T, ColMajor>> keep;
};
template <class T> struct ProdMM {
    T x;
    ConstMatrixView<T> a, b;
    std::shared_ptr<const Matrix<T, ColMajor>> keep_a, keep_b;
    ptrdiff_t colsize() const { return a.colsize(); }
    ptrdiff_t rowsize() const { return b.rowsize(); }
};

template <class T, int S> Matrix<T, S>::Matrix(const ProdMM<T>& p) : Matrix(p.colsize(), p.rowsize())
{ detail::multmm<false, T>(p.x, p.a, p.b, view()); }
template <class T, int S> Matrix<T, S>::Matrix(const ProdXM<T>& p) : Matrix(p.m)
{ *this *= p.x; }
template <class T, int S> Matrix<T, S>& Matrix<T, S>::operator=(const ProdMM<T>& p)
{
    if (p.colsize() != m_ || p.rowsize() != n_) {
        Matrix t(p);
        return *this = t;
    }
    detail::multmm<false, T>(p.x, p.a, p.b, view());
    resetDiv();
    return *this;
}
template <class T, int S> Matrix<T, S>& Matrix<T, S>::operator+=(const ProdMM<T>& p)
{ detail::multmm<true, T>(p.x, p.a, p.b, view()); resetDiv(); return *this; }
template <class T, int S> Matrix<T, S>& Matrix<T, S>::operator-=(const ProdMM<T>& p)
{ detail::multmm<true, T>(-p.x, p.a, p.b, view()); resetDiv(); return *this; }
template <class T, int S> Matrix<T, S>& Matrix<T, S>::operator*=(const ConstMatrixView<T>& m)
{
    // m *= m2 is m = m * m2 (examples/Matrix.cpp:336): square m2 keeps the
    // shape; C aliasing A is safe, the operands are staged first.
    if (m.colsize() != n_ || m.rowsize() != n_) throw Error("m *= m2: size mismatch");
    detail::multmm<false, T>(T(1), view(), m, view());
    resetDiv();
    return *this;
}

template <class T> const MatrixView<T>& MatrixView<T>::operator=(const ProdMM<T>& p) const
{ detail::multmm<false, T>(p.x, p.a, p.b, *this); return *this; }
template <class T> const MatrixView<T>& MatrixView<T>::operator+=(const ProdMM<T>& p) const
{ detail::multmm<true, T>(p.x, p.a, p.b, *this); return *this; }
template <class T> const MatrixView<T>& MatrixView<T>::operator-=(const ProdMM<T>& p) const
{ detail::multmm<true, T>(-p.x, p.a, p.b, *this); return *this; }
template <class T> const MatrixView<T>& MatrixView<T>::operator*=(const ConstMatrixView<T>& m) const
{
    if (m.colsize() != this->rowsize() || m.rowsize() != this->rowsize())
        throw Error("m *= m2: size mismatch");
    detail::multmm<false, T>(T(1), *this, m, *this);
    return *this;
}

// Operands of * : matrices, views, x*A.
template <class T, int SA, int SB>
ProdMM<T> operator*(const Matrix<T, SA>& A, const Matrix<T, SB>& B)
{ return ProdMM<T>{T(1), A.view(), B.view(), nullptr, nullptr}; }
template <class T, int SA>
ProdMM<T> operator*(const Matrix<T, SA>& A, const ConstMatrixView<T>& B)
{ return ProdMM<T>{T(1), A.view(), B, nullptr, nullptr}; }
template <class T, int SB>
ProdMM<T> operator*(const ConstMatrixView<T>& A, const Matrix<T, SB>& B)
{ return ProdMM<T>{T(1), A, B.view(), nullptr, nullptr}; }
template <class T>
ProdMM<T> operator*(const ConstMatrixView<T>& A, const ConstMatrixView<T>& B)
{ return ProdMM<T>{T(1), A, B, nullptr, nullptr}; }
template <class T>
ProdMM<T> operator*(const MatrixView<T>& A, const MatrixView<T>& B)
{ return ProdMM<T>{T(1), A, B, nullptr, nullptr}; }
template <class T, int SB>
ProdMM<T> operator*(const MatrixView<T>& A, const Matrix<T, SB>& B)
{ return ProdMM<T>{T(1), A, B.view(), nullptr, nullptr}; }
template <class T, int SA>
ProdMM<T> operator*(const Matrix<T, SA>& A, const MatrixView<T>& B)
{ return ProdMM<T>{T(1), A.view(), B, nullptr, nullptr}; }

template <class T, int S>
ProdXM<T> operator*(T x, const Matrix<T, S>& A) { return ProdXM<T>{x, A.view(), nullptr}; }
template <class T, int S>
ProdXM<T> operator*(const Matrix<T, S>& A, T x) { return ProdXM<T>{x, A.view(), nullptr}; }
template <class T>
ProdXM<T> operator*(T x, const ConstMatrixView<T>& A) { return ProdXM<T>{x, A, nullptr}; }
template <class T>
ProdXM<T> operator*(T x, const MatrixView<T>& A) { return ProdXM<T>{x, A, nullptr}; }
template <class T, int S>
ProdMM<T> operator*(const ProdXM<T>& xa, const Matrix<T, S>& B)
{ return ProdMM<T>{xa.x, xa.m, B.view(), xa.keep, nullptr}; }
template <class T>
ProdMM<T> operator*(const ProdXM<T>& xa, const ConstMatrixView<T>& B)
{ return ProdMM<T>{xa.x, xa.m, B, xa.keep, nullptr}; }
template <class T, int S>
ProdMM<T> operator*(const Matrix<T, S>& A, const ProdXM<T>& xb)
{ return ProdMM<T>{xb.x, A.view(), xb.m, nullptr, xb.keep}; }
template <class T> ProdMM<T> operator*(T x, ProdMM<T> p) { p.x = x * p.x; return p; }
template <class T> ProdMM<T> operator*(ProdMM<T> p, T x) { p.x = p.x * x; return p; }
template <class T> ProdMM<T> operator-(ProdMM<T> p) { p.x = -p.x; return p; }
// (A*B)*C and A*(B*C): the inner product is evaluated and kept alive
template <class T, int S>
ProdMM<T> operator*(const ProdMM<T>& p, const Matrix<T, S>& C)
{
    auto t = std::make_shared<const Matrix<T, ColMajor>>(p);
    return ProdMM<T>{T(1), t->view(), C.view(), t, nullptr};
}
template <class T, int S>
ProdMM<T> operator*(const Matrix<T, S>& A, const ProdMM<T>& p)
{
    auto t = std::make_shared<const Matrix<T, ColMajor>>(p);
    return ProdMM<T>{T(1), A.view(), t->view(), nullptr, t};
}

// C += x A B without a temporary: AddMM(C, A, B, x) (kept for callers of
// earlier versions of this header; C += x*A*B is the TMV spelling).
template <class T, int SA, int SB, int SC>
void AddMM(Matrix<T, SC>& C, const Matrix<T, SA>& A, const Matrix<T, SB>& B, T x = T(1))
{ MultMM<true>(x, A, B, C); }

// ---------------------------------------------------------------------------
// Division: x = b/m, X = B/m, x = b%m (QuotVM / QuotMM, TMV_MatrixArith.h)
// ---------------------------------------------------------------------------
template <class T, int S>
Vector<T> operator/(const Vector<T>& b, const Matrix<T, S>& m)
{
    if (b.size() != m.colsize()) throw Error("b/m: size mismatch");
    if (m.usesQR()) {
        Vector<T> x(m.rowsize());
        m.qrd().lDivCM(b.cptr(), x.ptr(), 1);
        m.doneDiv();
        return x;
    }
    Vector<T> x = b;
    m.lud().lDivEqCM(x.ptr(), 1);
    m.doneDiv();
    return x;
}

template <class T, int S>
Vector<T> operator%(const Vector<T>& b, const Matrix<T, S>& m)
{
    if (b.size() != m.rowsize()) throw Error("b%m: size mismatch");
    Vector<T> x = b;
    m.lud().rDivEqCM(x.ptr(), 1);
    m.doneDiv();
    return x;
}

template <class T, int SB, int S>
Matrix<T, ColMajor> operator/(const Matrix<T, SB>& B, const Matrix<T, S>& m)
{
    if (B.colsize() != m.colsize()) throw Error("B/m: size mismatch");
    if (m.usesQR()) {
        const Matrix<T, ColMajor> Bc(B);
        Matrix<T, ColMajor> X(m.rowsize(), B.rowsize());
        m.qrd().lDivCM(Bc.cptr(), X.ptr(), B.rowsize());
        m.doneDiv();
        return X;
    }
    Matrix<T, ColMajor> X(B);
    m.lud().lDivEqCM(X.ptr(), X.rowsize());
    m.doneDiv();
    return X;
}

// LU_Decompose(A, P): in place, A := packed L\U, P := interchanges
// (include/tmv/TMV_LUD.h:79-88, src/TMV_LUDecompose.cpp:514-538).  A
// conjugated view factorises its storage (the pivots of conj(A) are those
// of A); a column-major view goes to the GPU as it is; a row-major view
// (TMV transposes it in place, factorises and transposes back) and any
// other view are factorised through a column-major copy.  Either way A
// ends up holding the factors of the matrix it shows.  An m x n view takes
// min(m, n) pivot steps (P has m entries as in TMV's Permutation overload,
// include/tmv/TMV_Permutation.h:832-839, those past min(m, n) no
// interchange); the columns past m of a wide view hold L^-1 P A, as
// TMV's dgetrf seam gives them (real types; complex views are square).
template <class T>
void LU_Decompose(const MatrixView<T>& A, Permutation& P)
{
    if (A.isconj()) { LU_Decompose(A.conjugate(), P); return; }
    const ptrdiff_t m = A.colsize(), n = A.rowsize(), r = m < n ? m : n;
    P.resize(m);
    for (ptrdiff_t i = 0; i < m; ++i) P.getMem()[i] = i;
    if (r == 0) return;
    if (!A.iscm()) {
        Matrix<T, ColMajor> Ac(A);
        LU_Decompose(Ac.view(), P);
        A = Ac.view();
        return;
    }
    const ptrdiff_t ld = n > 1 ? A.stepj() : m;
    detail::DevBuf<T> a((size_t)(m * n));
    detail::DevBuf<int64_t> piv((size_t)r);
    a.upload2d(A.cptr(), ld, m, n);
    int64_t info = 0;
    if (m == n)
        detail::check(detail::Abi<T>::getrf(n, a.get(), n, piv.get(), &info), "LU_Decompose");
    else
        detail::check(detail::Abi<T>::getrf_mn(m, n, a.get(), m, piv.get(), &info), "LU_Decompose");
    a.download2d(A.ptr(), ld, m, n);
    std::vector<int64_t> hp((size_t)r);
    piv.download(hp.data(), (size_t)r);
    for (ptrdiff_t i = 0; i < r; ++i) P.getMem()[i] = (ptrdiff_t)hp[(size_t)i];
}

template <class T, int S>
void LU_Decompose(Matrix<T, S>& A, Permutation& P)
{
    LU_Decompose(A.view(), P);
    A.resetDiv();
}

// Mixed-precision x = b/m (FP32 factorisation + FP64 refinement to a
// backward error <= tol).  Not in TMV: an extension of the same division.
inline Vector<double> SolveMixed(const Vector<double>& b, const Matrix<double, ColMajor>& m,
                                 double tol = 1e-12, tmvx_ir_info* out = nullptr)
{
    const ptrdiff_t n = m.colsize();
    if (!m.isSquare() || b.size() != n) throw Error("SolveMixed: size mismatch");
    Vector<double> x(n);
    if (n == 0) return x;
    detail::DevBuf<double> a((size_t)(n * n)), bb((size_t)n), xx((size_t)n);
    a.upload(m.cptr(), (size_t)(n * n));
    bb.upload(b.cptr(), (size_t)n);
    tmvx_ir_info info;
    detail::check(tmvx_gesv_mixed_f64(TMVX_LOWPREC_F32, n, 1, a.get(), n, bb.get(), n, xx.get(), n,
                                      tol, 50, &info, nullptr), "SolveMixed");
    xx.download(x.ptr(), (size_t)n);
    if (out) *out = info;
    return x;
}

} // namespace tmv

#endif // TMVX_TMV_H
