// Inverse and determinant from LU factors.
//
// Counterparts of TMV's LU_Inverse (src/TMV_LUInverse.cpp:236-300; the
// LUDiv makeInverse of m.inverse()) and LUDiv::det / logDet
// (src/TMV_LUD.cpp:133-156).  The inverse is X = A^-1 I: the identity's
// rows permuted by the interchange list, then unit-lower L and U on all n
// columns at once (the blocked multi-RHS triangular solves, MFMA GEMM
// updates).  logDet sums log|U_ii| in FP64 over one workgroup; the sign is
// the permutation's parity times the signs of the diagonal; det = sign *
// exp(logDet), as TMV forms it.  A zero diagonal gives sign 0 and det 0.
#include "tmvx_lu.h"

#include <cmath>

namespace tmvx {

template <typename T>
__global__ void identity_kernel(i64 n, T* __restrict__ X, i64 ldx)
{
    const i64 j = blockIdx.y;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        X[i + j * ldx] = i == j ? T(1) : T(0);
}

// out[0] = sum log|U_ii|, out[1] = sign (+1, -1, or 0 if a diagonal is 0).
template <typename T>
__global__ void __launch_bounds__(1024)
logdet_kernel(i64 n, const T* __restrict__ LU, i64 lda, const i64* __restrict__ piv,
              double* __restrict__ out)
{
    __shared__ double s_sum[16];
    __shared__ int s_neg[16];
    __shared__ int s_zero[16];
    double sum = 0.0;
    int neg = 0, zero = 0;
    for (i64 i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = (double)LU[i + i * lda];
        if (d == 0.0) zero = 1;
        else {
            sum += log(fabs(d));
            if (d < 0.0) neg ^= 1;
        }
        if (piv[i] != i) neg ^= 1;                 // one interchange
    }
    // reduce within the wave, then across waves
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_down(sum, off);
        neg ^= __shfl_down(neg, off);
        zero |= __shfl_down(zero, off);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_sum[wave] = sum; s_neg[wave] = neg; s_zero[wave] = zero; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        int tneg = 0, tzero = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            tot += s_sum[w];
            tneg ^= s_neg[w];
            tzero |= s_zero[w];
        }
        out[0] = tzero ? -INFINITY : tot;
        out[1] = tzero ? 0.0 : (tneg ? -1.0 : 1.0);
    }
}

// out[0] = min_i |U_ii|, out[1] = max_i |U_ii|, with |z| = |re| + |im| for
// complex entries (CPLX: T is the real type, entries interleaved).
template <typename T, bool CPLX>
__global__ void __launch_bounds__(1024)
diag_range_kernel(i64 n, const T* __restrict__ LU, i64 lda, double* __restrict__ out)
{
    __shared__ double s_mn[16], s_mx[16];
    double mn = INFINITY, mx = 0.0;
    for (i64 i = threadIdx.x; i < n; i += blockDim.x) {
        double a;
        if constexpr (CPLX) {
            const T* e = LU + 2 * (i + i * lda);
            a = fabs((double)e[0]) + fabs((double)e[1]);
        } else {
            a = fabs((double)LU[i + i * lda]);
        }
        mn = fmin(mn, a);
        mx = fmax(mx, a);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_down(mn, off));
        mx = fmax(mx, __shfl_down(mx, off));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { s_mn[wave] = mn; s_mx[wave] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            mn = fmin(mn, s_mn[w]);
            mx = fmax(mx, s_mx[w]);
        }
        out[0] = mn;
        out[1] = mx;
    }
}

template <typename T, bool CPLX>
int diag_range_impl(i64 n, const T* LU, i64 lda, double* dmin, double* dmax, hipStream_t s)
{
    if (n < 0 || lda < (n > 0 ? n : 1) || !dmin || !dmax) return TMVX_ERR_ARG;
    if (n == 0) { *dmin = INFINITY; *dmax = 0.0; return TMVX_OK; }
    if (!LU) return TMVX_ERR_ARG;
    double* d = nullptr;
    if (hipMalloc((void**)&d, 2 * sizeof(double)) != hipSuccess) return TMVX_ERR_ALLOC;
    hipLaunchKernelGGL((diag_range_kernel<T, CPLX>), dim3(1), dim3(1024), 0, s, n, LU, lda, d);
    int rc = hip_status(hipGetLastError());
    double h[2] = {0.0, 0.0};
    if (rc == TMVX_OK &&
        hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    hipFree(d);
    *dmin = h[0];
    *dmax = h[1];
    return rc;
}

template <typename T>
int getri_impl(i64 n, const T* LU, i64 lda, const i64* piv, T* X, i64 ldx, hipStream_t s)
{
    if (n < 0 || lda < (n > 0 ? n : 1) || ldx < (n > 0 ? n : 1)) return TMVX_ERR_ARG;
    if (n == 0) return TMVX_OK;
    if (!LU || !piv || !X) return TMVX_ERR_ARG;
    if (n > 65535) return TMVX_ERR_UNSUPPORTED;            // grid.y limit of the fill
    hipLaunchKernelGGL(identity_kernel<T>, dim3((unsigned)((n + 255) / 256 < 64 ? (n + 255) / 256 : 64),
                                                (unsigned)n),
                       dim3(256), 0, s, n, X, ldx);
    TMVX_CHECK_LAUNCH();
    int rc;
    if constexpr (sizeof(T) == 8)
        rc = tmvx_getrs_f64(0, n, n, LU, lda, piv, X, ldx, 1, s);
    else
        rc = tmvx_getrs_f32(0, n, n, LU, lda, piv, X, ldx, 1, s);
    return rc;
}

template <typename T>
int logdet_impl(i64 n, const T* LU, i64 lda, const i64* piv, double* logdet, double* sign,
                hipStream_t s)
{
    if (n < 0 || lda < (n > 0 ? n : 1) || !logdet || !sign) return TMVX_ERR_ARG;
    if (n == 0) { *logdet = 0.0; *sign = 1.0; return TMVX_OK; }
    if (!LU || !piv) return TMVX_ERR_ARG;
    double* d = nullptr;
    if (hipMalloc((void**)&d, 2 * sizeof(double)) != hipSuccess) return TMVX_ERR_ALLOC;
    hipLaunchKernelGGL(logdet_kernel<T>, dim3(1), dim3(1024), 0, s, n, LU, lda, piv, d);
    int rc = hip_status(hipGetLastError());
    double h[2] = {0.0, 0.0};
    if (rc == TMVX_OK &&
        hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    hipFree(d);
    *logdet = h[0];
    *sign = h[1];
    return rc;
}

} // namespace tmvx

using namespace tmvx;

extern "C" {

int tmvx_getri_f64(int64_t n, const double* LU, int64_t lda, const int64_t* piv, double* X,
                   int64_t ldx, void* stream)
{ return getri_impl<double>(n, LU, lda, piv, X, ldx, (hipStream_t)stream); }

int tmvx_getri_f32(int64_t n, const float* LU, int64_t lda, const int64_t* piv, float* X,
                   int64_t ldx, void* stream)
{ return getri_impl<float>(n, LU, lda, piv, X, ldx, (hipStream_t)stream); }

int tmvx_logdet_f64(int64_t n, const double* LU, int64_t lda, const int64_t* piv,
                    double* logdet, double* sign, void* stream)
{ return logdet_impl<double>(n, LU, lda, piv, logdet, sign, (hipStream_t)stream); }

int tmvx_logdet_f32(int64_t n, const float* LU, int64_t lda, const int64_t* piv,
                    double* logdet, double* sign, void* stream)
{ return logdet_impl<float>(n, LU, lda, piv, logdet, sign, (hipStream_t)stream); }

int tmvx_lu_diag_range_f64(int64_t n, const double* LU, int64_t lda, double* dmin, double* dmax,
                           void* stream)
{ return diag_range_impl<double, false>(n, LU, lda, dmin, dmax, (hipStream_t)stream); }

int tmvx_lu_diag_range_f32(int64_t n, const float* LU, int64_t lda, double* dmin, double* dmax,
                           void* stream)
{ return diag_range_impl<float, false>(n, LU, lda, dmin, dmax, (hipStream_t)stream); }

int tmvx_lu_diag_range_z(int64_t n, const double* LU, int64_t lda, double* dmin, double* dmax,
                         void* stream)
{ return diag_range_impl<double, true>(n, LU, lda, dmin, dmax, (hipStream_t)stream); }

int tmvx_lu_diag_range_c(int64_t n, const float* LU, int64_t lda, double* dmin, double* dmax,
                         void* stream)
{ return diag_range_impl<float, true>(n, LU, lda, dmin, dmax, (hipStream_t)stream); }

} // extern "C"
