// tmvx_tmv.h -- TMV-shaped C++ API over the tmvx C ABI (header only).
//
// The part of TMV 0.76's user surface that the dense GEMM + LU path serves,
// with TMV's names, argument meaning and error behaviour, running on the
// MI355X through include/tmvx.h:
//   tmv::Matrix<T, tmv::ColMajor|tmv::RowMajor>   (TMV_Matrix.h)
//   C = A*B, C += A*B, C -= A*B, MultMM<add>(alpha, A, B, C)
//                                                  (TMV_MatrixArith.h:700-812,
//                                                   TMV_MatrixArithFunc.h:53-57)
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
//   LU_Decompose(Matrix&, Permutation&)             (TMV_LUD.h:84-88)
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
#include <stdexcept>
#include <string>
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
private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

template <class T> struct Abi;
template <> struct Abi<double> {
    static int gemm(int la, int lb, int64_t m, int64_t n, int64_t k, double al, const double* A,
                    int64_t lda, const double* B, int64_t ldb, double be, double* C, int64_t ldc)
    { return tmvx_gemm_f64(la, lb, m, n, k, al, A, lda, B, ldb, be, C, ldc, nullptr); }
    static int getrf(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_f64(n, A, lda, piv, info, nullptr); }
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
    static int gemm(int la, int lb, int64_t m, int64_t n, int64_t k, float al, const float* A,
                    int64_t lda, const float* B, int64_t ldb, float be, float* C, int64_t ldc)
    { return tmvx_gemm_f32(la, lb, m, n, k, al, A, lda, B, ldb, be, C, ldc, nullptr); }
    static int getrf(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_f32(n, A, lda, piv, info, nullptr); }
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
    static int gemm(int la, int lb, int64_t m, int64_t n, int64_t k, Z al, const Z* A,
                    int64_t lda, const Z* B, int64_t ldb, Z be, Z* C, int64_t ldc)
    {
        return tmvx_gemm_z(la, lb, 0, 0, m, n, k, al.real(), al.imag(),
                           reinterpret_cast<const double*>(A), lda,
                           reinterpret_cast<const double*>(B), ldb, be.real(), be.imag(),
                           reinterpret_cast<double*>(C), ldc, nullptr);
    }
    static int getrf(int64_t n, Z* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_z(n, reinterpret_cast<double*>(A), lda, piv, info, nullptr); }
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
    static int gemm(int la, int lb, int64_t m, int64_t n, int64_t k, Z al, const Z* A,
                    int64_t lda, const Z* B, int64_t ldb, Z be, Z* C, int64_t ldc)
    {
        return tmvx_gemm_c(la, lb, 0, 0, m, n, k, al.real(), al.imag(),
                           reinterpret_cast<const float*>(A), lda,
                           reinterpret_cast<const float*>(B), ldb, be.real(), be.imag(),
                           reinterpret_cast<float*>(C), ldc, nullptr);
    }
    static int getrf(int64_t n, Z* A, int64_t lda, int64_t* piv, int64_t* info)
    { return tmvx_getrf_c(n, reinterpret_cast<float*>(A), lda, piv, info, nullptr); }
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
    explicit LUDiv(const Matrix<T, S>& A)
        : n_(A.colsize()), istrans_(A.isrm()), P_(A.colsize())
    {
        if (!A.isSquare()) throw Error("LU division requires a square matrix");
        lu_.resize((size_t)(n_ * n_));
        piv_.resize((size_t)n_);
        if (n_ == 0) return;
        // Column-major view of the storage: A itself, or A^T when row-major.
        lu_.upload(A.cptr(), (size_t)(n_ * n_));
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
    // min|U_ii| <= eps * max|U_ii| (src/TMV_LUD.cpp:204-209)
    bool isSingular() const
    {
        Matrix<T, ColMajor> M = getLU();
        T mn = std::numeric_limits<T>::infinity(), mx = 0;
        for (ptrdiff_t i = 0; i < n_; ++i) { mn = std::fmin(mn, std::fabs(M(i, i))); mx = std::fmax(mx, std::fabs(M(i, i))); }
        return mn <= std::numeric_limits<T>::epsilon() * mx;
    }

    // m := A^-1 m (LDivEq), nr right-hand sides stored column-major.
    void lDivEqCM(T* m, ptrdiff_t nr) const { solve(m, nr, false); }
    // m := m A^-1 for row vectors, i.e. A^T y = m (RDivEq).
    void rDivEqCM(T* m, ptrdiff_t nr) const { solve(m, nr, true); }

private:
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
template <bool add, class T, int SA, int SB, int SC>
void MultMM(T alpha, const Matrix<T, SA>& A, const Matrix<T, SB>& B, Matrix<T, SC>& C)
{
    if (A.colsize() != C.colsize() || A.rowsize() != B.colsize() || B.rowsize() != C.rowsize())
        throw Error("MultMM: size mismatch");
    const ptrdiff_t m = C.colsize(), n = C.rowsize(), k = A.rowsize();
    if (m == 0 || n == 0) return;
    detail::DevBuf<T> a((size_t)(m * k)), b((size_t)(k * n)), c((size_t)(m * n));
    if (m * k) a.upload(A.cptr(), (size_t)(m * k));
    if (k * n) b.upload(B.cptr(), (size_t)(k * n));
    // C row-major: compute C^T = B^T A^T into column-major storage
    // (src/TMV_MultMM.cpp:447-448).
    const bool crm = SC == RowMajor;
    if (add) c.upload(C.cptr(), (size_t)(m * n));
    const int la = SA == RowMajor ? TMVX_ROWMAJOR : TMVX_COLMAJOR;
    const int lb = SB == RowMajor ? TMVX_ROWMAJOR : TMVX_COLMAJOR;
    const int64_t lda = SA == RowMajor ? k : m, ldb = SB == RowMajor ? n : k;
    if (!crm) {
        detail::check(detail::Abi<T>::gemm(la, lb, m, n, k, alpha, a.get(), lda, b.get(), ldb,
                                           add ? T(1) : T(0), c.get(), m), "MultMM");
    } else {
        // C^T (n x m, column-major) = B^T A^T: the layouts of A and B flip.
        detail::check(detail::Abi<T>::gemm(lb ^ 1, la ^ 1, n, m, k, alpha, b.get(), ldb, a.get(),
                                           lda, add ? T(1) : T(0), c.get(), n), "MultMM");
    }
    c.download(C.ptr(), (size_t)(m * n));
}

template <class T, int SA, int SB>
Matrix<T, ColMajor> operator*(const Matrix<T, SA>& A, const Matrix<T, SB>& B)
{
    Matrix<T, ColMajor> C(A.colsize(), B.rowsize());
    MultMM<false>(T(1), A, B, C);
    return C;
}

// C += A*B / C -= A*B without a temporary: AddMM(C, A, B, x) = C += x A B.
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
// (include/tmv/TMV_LUD.h:84-88).  Row-major A is factorised as stored
// (LU of the column-major view), as TMV's LU_Decompose on a row-major
// matrix transposes in place and factorises the transpose.
template <class T>
void LU_Decompose(Matrix<T, ColMajor>& A, Permutation& P)
{
    if (!A.isSquare()) throw Error("LU_Decompose: square matrices only");
    const ptrdiff_t n = A.colsize();
    P.resize(n);
    if (n == 0) return;
    detail::DevBuf<T> a((size_t)(n * n));
    detail::DevBuf<int64_t> piv((size_t)n);
    a.upload(A.cptr(), (size_t)(n * n));
    int64_t info = 0;
    detail::check(detail::Abi<T>::getrf(n, a.get(), n, piv.get(), &info), "LU_Decompose");
    a.download(A.ptr(), (size_t)(n * n));
    std::vector<int64_t> hp((size_t)n);
    piv.download(hp.data(), (size_t)n);
    for (ptrdiff_t i = 0; i < n; ++i) P.getMem()[i] = (ptrdiff_t)hp[(size_t)i];
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
