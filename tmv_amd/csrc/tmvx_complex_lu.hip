// Complex LU with partial pivoting and the solves with its factors
// (interleaved std::complex storage).
//
// TMV's LU_Decompose and LU_LDivEq / LU_RDivEq for complex T
// (src/TMV_LUDecompose.cpp:217-407, src/TMV_LUDiv.cpp:185-226, 369-410): the
// pivot is the candidate of largest modulus (TMV_ABS), first index on ties
// (FindMaxAbsElement, src/TMV_Vector.cpp:473-504); a column whose largest
// candidate underflows is zeroed from the diagonal down and interchanged with
// itself (src/TMV_LUDecompose.cpp:162-169); the column below the pivot is
// scaled by the pivot's reciprocal.
//
// Blocked right-looking, 64-column panels:
//   panel: ONE workgroup factors the panel column by column (pivot search
//     by a block reduction, the interchange applied to the whole row, the
//     column scaled and the panel's remaining columns updated); a simple,
//     latency-bound form -- the complex path is not a measured workload;
//   U12 = L11^-1 A12: one thread per column;
//   A22 -= L21 U12: the complex GEMM (four real MFMA products).
// Solves: interchanges by one thread per right-hand side, then triangular
// column sweeps by one workgroup per right-hand side.
#include "tmvx_common.h"

#include <cfloat>

namespace tmvx {

template <typename R> struct CplxT;
template <> struct CplxT<double> { typedef double2 type; static constexpr double rmin = DBL_MIN; };
template <> struct CplxT<float> { typedef float2 type; static constexpr float rmin = FLT_MIN; };

__global__ void set_i64_kernel(i64* p, i64 v) { *p = v; }
__global__ void set_int_kernel(int* p, int v) { *p = v; }

template <typename Z> __device__ __forceinline__ Z cmul(Z a, Z b)
{
    Z r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename Z> __device__ __forceinline__ Z csub(Z a, Z b)
{
    Z r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
// a / b by Smith's method (scaled: no overflow for large |b|).
template <typename Z> __device__ __forceinline__ Z cdiv(Z a, Z b)
{
    Z r;
    if (fabs(b.x) >= fabs(b.y)) {
        const auto t = b.y / b.x, d = b.x + b.y * t;
        r.x = (a.x + a.y * t) / d;
        r.y = (a.y - a.x * t) / d;
    } else {
        const auto t = b.x / b.y, d = b.x * t + b.y;
        r.x = (a.x * t + a.y) / d;
        r.y = (a.y * t - a.x) / d;
    }
    return r;
}

constexpr int ZP_NB = 64;        // panel width
constexpr int ZP_NTH = 1024;     // panel workgroup

// Factor panel columns [j0, j1) of the n x n matrix; rows swapped over all n
// columns.  One workgroup.
template <typename R>
__global__ void __launch_bounds__(ZP_NTH)
zpanel_kernel(i64 n, typename CplxT<R>::type* __restrict__ A, i64 lda, i64 j0, i64 j1,
              i64* __restrict__ piv, i64* __restrict__ first_zero)
{
    typedef typename CplxT<R>::type Z;
    constexpr int NW = ZP_NTH / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    __shared__ R s_v[NW];
    __shared__ i64 s_i[NW];
    __shared__ i64 s_p;
    __shared__ int s_zero;
    __shared__ Z s_prow[ZP_NB];
    __shared__ Z s_inv;
    for (i64 j = j0; j < j1; ++j) {
        // pivot search: largest modulus, lowest row on ties
        R bv = R(-1);
        i64 bi = n;
        for (i64 i = j + t; i < n; i += ZP_NTH) {
            const Z a = A[i + j * lda];
            const R v = hypot(a.x, a.y);
            if (v > bv) { bv = v; bi = i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const R v2 = __shfl_xor(bv, off);
            const i64 i2 = __shfl_xor(bi, off);
            if (v2 > bv || (v2 == bv && i2 < bi)) { bv = v2; bi = i2; }
        }
        if (lane == 0) { s_v[wave] = bv; s_i[wave] = bi; }
        __syncthreads();
        if (t == 0) {
            R v = s_v[0];
            i64 ii = s_i[0];
            for (int q = 1; q < NW; ++q)
                if (s_v[q] > v || (s_v[q] == v && s_i[q] < ii)) { v = s_v[q]; ii = s_i[q]; }
            const bool zero = !(v >= CplxT<R>::rmin);
            s_zero = zero ? 1 : 0;
            s_p = zero ? j : ii;
            piv[j] = zero ? j : ii;
            if (zero)
                atomicMin(reinterpret_cast<unsigned long long*>(first_zero),
                          (unsigned long long)(j + 1));
        }
        __syncthreads();
        const i64 p = s_p;
        if (s_zero) {
            Z z0;
            z0.x = R(0);
            z0.y = R(0);
            for (i64 i = j + t; i < n; i += ZP_NTH) A[i + j * lda] = z0;
            __syncthreads();
            continue;
        }
        if (p != j)
            for (i64 c = t; c < n; c += ZP_NTH) {
                const Z x = A[j + c * lda];
                A[j + c * lda] = A[p + c * lda];
                A[p + c * lda] = x;
            }
        __syncthreads();
        if (t < ZP_NB && j + 1 + t < j1) s_prow[t] = A[j + (j + 1 + t) * lda];
        if (t == 0) {
            Z one;
            one.x = R(1);
            one.y = R(0);
            s_inv = cdiv(one, A[j + j * lda]);
        }
        __syncthreads();
        const Z inv = s_inv;
        const int nc = (int)(j1 - j - 1);
        for (i64 i = j + 1 + t; i < n; i += ZP_NTH) {
            const Z l = cmul(A[i + j * lda], inv);
            A[i + j * lda] = l;
            for (int c = 0; c < nc; ++c)
                A[i + (j + 1 + c) * lda] = csub(A[i + (j + 1 + c) * lda], cmul(l, s_prow[c]));
        }
        __syncthreads();
    }
}

// U12 := L11^-1 A12 for the panel rows [j0, j0 + jb) and columns [c0, n).
template <typename R>
__global__ void ztrsm_u12_kernel(i64 n, typename CplxT<R>::type* __restrict__ A, i64 lda, i64 j0,
                                 int jb, i64 c0)
{
    typedef typename CplxT<R>::type Z;
    const i64 c = c0 + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    Z* col = A + c * lda + j0;
    const Z* L = A + j0 + j0 * lda;
    for (int r = 1; r < jb; ++r) {
        Z acc = col[r];
        for (int q = 0; q < r; ++q) acc = csub(acc, cmul(L[r + (i64)q * lda], col[q]));
        col[r] = acc;
    }
}

template <typename R>
int zgetrf_impl(i64 n, R* Ar, i64 lda, i64* piv, i64* info, hipStream_t s)
{
    typedef typename CplxT<R>::type Z;
    if (n < 0 || lda < (n > 0 ? n : 1) || !info) return TMVX_ERR_ARG;
    *info = 0;
    if (n == 0) return TMVX_OK;
    if (!Ar || !piv) return TMVX_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(Ar) % sizeof(Z)) return TMVX_ERR_ARG;
    Z* A = reinterpret_cast<Z*>(Ar);
    i64* d_fz = nullptr;
    if (hipMallocAsync((void**)&d_fz, sizeof(i64), s) != hipSuccess) return TMVX_ERR_ALLOC;
    const i64 big = n + 1;
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, d_fz, big);
    int rc = hip_status(hipGetLastError());
    for (i64 j0 = 0; j0 < n && rc == TMVX_OK; j0 += ZP_NB) {
        const int jb = (int)(n - j0 < ZP_NB ? n - j0 : ZP_NB);
        const i64 j1 = j0 + jb;
        hipLaunchKernelGGL(zpanel_kernel<R>, dim3(1), dim3(ZP_NTH), 0, s, n, A, lda, j0, j1, piv,
                           d_fz);
        if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
        const i64 nr = n - j1;
        if (nr <= 0) break;
        hipLaunchKernelGGL(ztrsm_u12_kernel<R>, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, s,
                           n, A, lda, j0, jb, j1);
        if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
        rc = gemm_complex<R>(TMVX_COLMAJOR, TMVX_COLMAJOR, 0, 0, nr, nr, jb, R(-1), R(0),
                             reinterpret_cast<const R*>(A + j1 + j0 * lda), lda,
                             reinterpret_cast<const R*>(A + j0 + j1 * lda), lda, R(1), R(0),
                             reinterpret_cast<R*>(A + j1 + j1 * lda), lda, s);
    }
    i64 fz = big;
    if (hipMemcpyAsync(&fz, d_fz, sizeof(i64), hipMemcpyDeviceToHost, s) != hipSuccess &&
        rc == TMVX_OK)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(d_fz, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    *info = fz <= n ? fz : 0;
    return rc;
}

// Interchanges on B's rows, one thread per right-hand side: in order
// (forward) or in reverse order (undo).
template <typename R>
__global__ void zpermute_kernel(i64 n, i64 nrhs, typename CplxT<R>::type* __restrict__ B, i64 ldb,
                                const i64* __restrict__ piv, int reverse)
{
    typedef typename CplxT<R>::type Z;
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs) return;
    Z* b = B + r * ldb;
    for (i64 q = 0; q < n; ++q) {
        const i64 j = reverse ? n - 1 - q : q;
        const i64 p = piv[j];
        if (p != j && p >= 0 && p < n) {
            const Z x = b[j];
            b[j] = b[p];
            b[p] = x;
        }
    }
}

// Triangular column sweep on one right-hand side per workgroup.
//   mode 0: L x = b (unit lower), forward      mode 1: U x = b (upper), backward
//   mode 2: U^T x = b (lower), forward         mode 3: L^T x = b (unit upper), backward
// A zero U diagonal met by a non-zero entry sets *singular.
template <typename R>
__global__ void __launch_bounds__(256)
zsweep_kernel(i64 n, const typename CplxT<R>::type* __restrict__ LU, i64 lda,
              typename CplxT<R>::type* __restrict__ B, i64 ldb, int mode, int* __restrict__ singular)
{
    typedef typename CplxT<R>::type Z;
    __shared__ Z s_x;
    Z* b = B + (i64)blockIdx.x * ldb;
    const int t = threadIdx.x;
    const bool fwd = mode == 0 || mode == 2;
    const bool diag = mode == 1 || mode == 2;
    for (i64 q = 0; q < n; ++q) {
        const i64 j = fwd ? q : n - 1 - q;
        if (t == 0) {
            Z x = b[j];
            if (diag) {
                const Z d = LU[j + j * lda];
                if (d.x == R(0) && d.y == R(0)) {
                    if (x.x != R(0) || x.y != R(0)) *singular = 1;
                    x.x = R(0);
                    x.y = R(0);
                } else {
                    x = cdiv(x, d);
                }
                b[j] = x;
            }
            s_x = x;
        }
        __syncthreads();
        const Z x = s_x;
        if (x.x != R(0) || x.y != R(0)) {
            if (fwd) {
                for (i64 i = j + 1 + t; i < n; i += blockDim.x) {
                    // mode 0: L(i, j); mode 2: U^T(i, j) = U(j, i)
                    const Z a = mode == 0 ? LU[i + j * lda] : LU[j + i * lda];
                    b[i] = csub(b[i], cmul(a, x));
                }
            } else {
                for (i64 i = t; i < j; i += blockDim.x) {
                    // mode 1: U(i, j); mode 3: L^T(i, j) = L(j, i)
                    const Z a = mode == 1 ? LU[i + j * lda] : LU[j + i * lda];
                    b[i] = csub(b[i], cmul(a, x));
                }
            }
        }
        __syncthreads();
    }
}

template <typename R>
int zgetrs_impl(int trans, i64 n, i64 nrhs, const R* LUr, i64 lda, const i64* piv, R* Br, i64 ldb,
                hipStream_t s)
{
    typedef typename CplxT<R>::type Z;
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1)) return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!LUr || !piv || !Br) return TMVX_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(LUr) % sizeof(Z) || reinterpret_cast<uintptr_t>(Br) % sizeof(Z))
        return TMVX_ERR_ARG;
    if (nrhs > 0x7fffffff) return TMVX_ERR_UNSUPPORTED;
    const Z* LU = reinterpret_cast<const Z*>(LUr);
    Z* B = reinterpret_cast<Z*>(Br);
    int* d_sing = nullptr;
    if (hipMallocAsync((void**)&d_sing, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, d_sing, 0);
    int rc = hip_status(hipGetLastError());
    const dim3 gp((unsigned)((nrhs + 63) / 64)), bp(64), gs((unsigned)nrhs), bs(256);
    if (rc == TMVX_OK && !trans) {
        hipLaunchKernelGGL(zpermute_kernel<R>, gp, bp, 0, s, n, nrhs, B, ldb, piv, 0);
        hipLaunchKernelGGL(zsweep_kernel<R>, gs, bs, 0, s, n, LU, lda, B, ldb, 0, d_sing);
        hipLaunchKernelGGL(zsweep_kernel<R>, gs, bs, 0, s, n, LU, lda, B, ldb, 1, d_sing);
        rc = hip_status(hipGetLastError());
    } else if (rc == TMVX_OK) {
        hipLaunchKernelGGL(zsweep_kernel<R>, gs, bs, 0, s, n, LU, lda, B, ldb, 2, d_sing);
        hipLaunchKernelGGL(zsweep_kernel<R>, gs, bs, 0, s, n, LU, lda, B, ldb, 3, d_sing);
        hipLaunchKernelGGL(zpermute_kernel<R>, gp, bp, 0, s, n, nrhs, B, ldb, piv, 1);
        rc = hip_status(hipGetLastError());
    }
    int h_sing = 0;
    if (hipMemcpyAsync(&h_sing, d_sing, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess &&
        rc == TMVX_OK)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(d_sing, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    if (rc == TMVX_OK && h_sing) rc = TMVX_ERR_SINGULAR;
    return rc;
}

} // namespace tmvx

using namespace tmvx;

extern "C" {

int tmvx_getrf_z(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return zgetrf_impl<double>(n, A, lda, piv, info, (hipStream_t)stream); }

int tmvx_getrf_c(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return zgetrf_impl<float>(n, A, lda, piv, info, (hipStream_t)stream); }

int tmvx_getrs_z(int trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                 const int64_t* piv, double* B, int64_t ldb, void* stream)
{ return zgetrs_impl<double>(trans, n, nrhs, LU, lda, piv, B, ldb, (hipStream_t)stream); }

int tmvx_getrs_c(int trans, int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                 const int64_t* piv, float* B, int64_t ldb, void* stream)
{ return zgetrs_impl<float>(trans, n, nrhs, LU, lda, piv, B, ldb, (hipStream_t)stream); }

} // extern "C"

// ---------------------------------------------------------------------------
// Determinant and inverse from the complex factors.
// ---------------------------------------------------------------------------
namespace tmvx {

// out[0] = sum log|U_ii| (FP64), out[1..2] = det(P) * prod U_ii / |U_ii|
// (0 with out[0] = -inf when a diagonal is 0): LUDiv::logDet for complex T
// (src/TMV_LUD.cpp:145-156; DiagMatrix::logDet multiplies the signs z/|z|,
// src/TMV_DiagMatrix.cpp:59-90).  One workgroup of 256 threads.
template <typename R>
__global__ void __launch_bounds__(256)
zlogdet_kernel(i64 n, const typename CplxT<R>::type* __restrict__ LU, i64 lda,
               const i64* __restrict__ piv, double* __restrict__ out)
{
    __shared__ double s_l[256], s_re[256], s_im[256];
    __shared__ int s_neg[256], s_zero[256];
    const int t = threadIdx.x;
    double sum = 0.0, pr = 1.0, pi = 0.0;
    int neg = 0, zero = 0;
    for (i64 i = t; i < n; i += blockDim.x) {
        const double x = (double)LU[i + i * lda].x, y = (double)LU[i + i * lda].y;
        const double a = hypot(x, y);
        if (a == 0.0) zero = 1;
        else {
            sum += log(a);
            const double ur = x / a, ui = y / a;
            const double nr = pr * ur - pi * ui;
            pi = pr * ui + pi * ur;
            pr = nr;
        }
        if (piv[i] != i) neg ^= 1;
    }
    s_l[t] = sum; s_re[t] = pr; s_im[t] = pi; s_neg[t] = neg; s_zero[t] = zero;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            s_l[t] += s_l[t + off];
            const double ar = s_re[t], ai = s_im[t], br = s_re[t + off], bi = s_im[t + off];
            s_re[t] = ar * br - ai * bi;
            s_im[t] = ar * bi + ai * br;
            s_neg[t] ^= s_neg[t + off];
            s_zero[t] |= s_zero[t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double sg = s_neg[0] ? -1.0 : 1.0;
        out[0] = s_zero[0] ? -INFINITY : s_l[0];
        out[1] = s_zero[0] ? 0.0 : sg * s_re[0];
        out[2] = s_zero[0] ? 0.0 : sg * s_im[0];
    }
}

template <typename R>
__global__ void zidentity_kernel(i64 n, typename CplxT<R>::type* __restrict__ X, i64 ldx)
{
    const i64 j = blockIdx.y;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        typename CplxT<R>::type z;
        z.x = i == j ? R(1) : R(0);
        z.y = R(0);
        X[i + j * ldx] = z;
    }
}

template <typename R>
int zlogdet_impl(i64 n, const R* LUr, i64 lda, const i64* piv, double* logdet, double* sre,
                 double* sim, hipStream_t s)
{
    typedef typename CplxT<R>::type Z;
    if (n < 0 || lda < (n > 0 ? n : 1) || !logdet || !sre || !sim) return TMVX_ERR_ARG;
    if (n == 0) { *logdet = 0.0; *sre = 1.0; *sim = 0.0; return TMVX_OK; }
    if (!LUr || !piv) return TMVX_ERR_ARG;
    if (reinterpret_cast<uintptr_t>(LUr) % sizeof(Z)) return TMVX_ERR_ARG;
    double* d = nullptr;
    if (hipMalloc((void**)&d, 3 * sizeof(double)) != hipSuccess) return TMVX_ERR_ALLOC;
    hipLaunchKernelGGL(zlogdet_kernel<R>, dim3(1), dim3(256), 0, s, n,
                       reinterpret_cast<const Z*>(LUr), lda, piv, d);
    int rc = hip_status(hipGetLastError());
    double h[3] = {0.0, 0.0, 0.0};
    if (rc == TMVX_OK && hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    hipFree(d);
    *logdet = h[0];
    *sre = h[1];
    *sim = h[2];
    return rc;
}

template <typename R>
int zgetri_impl(i64 n, const R* LU, i64 lda, const i64* piv, R* Xr, i64 ldx, hipStream_t s)
{
    typedef typename CplxT<R>::type Z;
    if (n < 0 || lda < (n > 0 ? n : 1) || ldx < (n > 0 ? n : 1)) return TMVX_ERR_ARG;
    if (n == 0) return TMVX_OK;
    if (!LU || !piv || !Xr) return TMVX_ERR_ARG;
    if (n > 65535) return TMVX_ERR_UNSUPPORTED;
    if (reinterpret_cast<uintptr_t>(Xr) % sizeof(Z)) return TMVX_ERR_ARG;
    const i64 gx = (n + 255) / 256;
    hipLaunchKernelGGL(zidentity_kernel<R>, dim3((unsigned)(gx < 64 ? gx : 64), (unsigned)n),
                       dim3(256), 0, s, n, reinterpret_cast<Z*>(Xr), ldx);
    TMVX_CHECK_LAUNCH();
    return zgetrs_impl<R>(0, n, n, LU, lda, piv, Xr, ldx, s);
}

} // namespace tmvx

extern "C" {

int tmvx_logdet_z(int64_t n, const double* LU, int64_t lda, const int64_t* piv, double* logdet,
                  double* sign_re, double* sign_im, void* stream)
{ return tmvx::zlogdet_impl<double>(n, LU, lda, piv, logdet, sign_re, sign_im, (hipStream_t)stream); }

int tmvx_getri_z(int64_t n, const double* LU, int64_t lda, const int64_t* piv, double* X,
                 int64_t ldx, void* stream)
{ return tmvx::zgetri_impl<double>(n, LU, lda, piv, X, ldx, (hipStream_t)stream); }

} // extern "C"
