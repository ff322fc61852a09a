// Adjoint of the LU solve X = A^-1 B, and a batched factorisation entry
// point for batches that include systems above the register-resident
// limit.
//
// Given the cotangent G = dL/dX of a scalar loss L:
//   gradB = A^-T G           (transposed solve with the stored LU factors)
//   gradA = -gradB X^T       (rank-nrhs update)
// TMV has no derivative code; the transposed solve is TMV's LU_RDivEq
// (x % A, src/TMV_LUDiv.cpp:369-410) and the rank-k update its MultMM
// (src/TMV_MultMM.cpp:404-477).
//
// Batched kernels: one wave per system solves A^T gB = G column by column
// with the stored factors (dot-product sweeps over contiguous columns of U
// and L); then many workgroups per system write gradA as
// -sum_c gB(:,c) X(k,c), 16 columns each (gradA, the n^2 output, is
// written once and read only when accumulating).  The recompute variant
// factors copies of the original matrices first, for callers that did not
// keep the factors.
#include "tmvx_lu.h"
#include "tmvx_pivot.h"
#include "tmvx_wave_solve.h"
#include "../../include/tmvx_batched.h"

#include <algorithm>
#include <climits>
#include <vector>

namespace tmvx {

constexpr int ADJ_RHS = 4;     // right-hand sides held in registers at once

template <typename T>
__global__ void __launch_bounds__(64)
adjoint_batched_kernel(int count, const int* __restrict__ nv, const int* __restrict__ nrhsv,
                       const T* const* __restrict__ LUv, const i64* const* __restrict__ pivv,
                       const T* const* __restrict__ Gv, T* const* __restrict__ gBv)
{
    const int lane = threadIdx.x & 63;
    for (int sys = blockIdx.x; sys < count; sys += gridDim.x) {
        const int n = nv[sys], nr = nrhsv[sys];
        if (n <= 0 || n > WS_NMAX || nr <= 0) continue;
        const T* LU = LUv[sys];
        const i64* piv = pivv[sys];
        const T* G = Gv[sys];
        T* gB = gBv[sys];
        // gradB = A^-T G, one right-hand side at a time (a runtime loop: the
        // solve is inlined once, its vector stays in registers)
        for (int c = 0; c < nr; ++c) {
            const i64 off = (i64)c * n;
            T x[WS_PR];
            wave_load(x, G + off, n, lane);
            wave_getrs<T, true>(1, n, LU, (i64)n, piv, x, lane);
            wave_store(x, gB + off, n, lane);
        }
    }
}

// gradA (+)= -gradB X^T for the systems the wave solves handled, spread
// over many workgroups (one wave per system writing its n^2 gradient left
// most of HBM idle): workgroup item (sys, kb) forms columns
// [ADJ_KC kb, ADJ_KC kb + ADJ_KC) of gradA, one row per thread.  The
// right-hand sides are taken in blocks of ADJ_RHS, the first block's sum
// written (or added, when accumulating) and each later block's added, in
// that order.
constexpr int ADJ_KC = 16;
template <typename T>
__global__ void __launch_bounds__(256)
adjoint_grad_kernel(int count, const int* __restrict__ nv, const int* __restrict__ nrhsv,
                    const T* const* __restrict__ Xv, const T* const* __restrict__ gBv,
                    T* const* __restrict__ gAv, int accumulate)
{
    constexpr int NKB = WS_NMAX / ADJ_KC;
    const i64 items = (i64)count * NKB;
    for (i64 item = blockIdx.x; item < items; item += gridDim.x) {
        const int sys = (int)(item / NKB), kb = (int)(item % NKB);
        const int n = nv[sys], nr = nrhsv[sys];
        if (n <= 0 || n > WS_NMAX || nr <= 0 || kb * ADJ_KC >= n) continue;
        const T* X = Xv[sys];
        const T* gB = gBv[sys];
        T* gA = gAv[sys];
        const int k1 = (kb + 1) * ADJ_KC < n ? (kb + 1) * ADJ_KC : n;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            for (int k = kb * ADJ_KC; k < k1; ++k) {
                T cur = T(0);
                for (int c0 = 0; c0 < nr; c0 += ADJ_RHS) {
                    const int cn = (nr - c0) < ADJ_RHS ? (nr - c0) : ADJ_RHS;
                    T v = T(0);
                    for (int c = 0; c < cn; ++c)
                        v -= gB[i + (i64)(c0 + c) * n] * X[k + (i64)(c0 + c) * n];
                    if (c0 == 0) cur = accumulate ? gA[i + (i64)k * n] + v : v;
                    else cur = cur + v;
                }
                gA[i + (i64)k * n] = cur;
            }
        }
    }
}

__global__ void batched_copy_kernel(int count, const int* __restrict__ nv,
                                    const char* const* __restrict__ src,
                                    char* const* __restrict__ dst, int esz)
{
    for (int sys = blockIdx.x; sys < count; sys += gridDim.x) {
        const i64 bytes = (i64)nv[sys] * nv[sys] * esz;
        const char* a = src[sys];
        char* b = dst[sys];
        for (i64 i = threadIdx.x; i < bytes; i += blockDim.x) b[i] = a[i];
    }
}

__global__ void scatter_info_kernel(int cnt, const int* __restrict__ idx,
                                    const int32_t* __restrict__ src, int32_t* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) dst[idx[i]] = src[i];
}

template <typename T>
static int adjoint_launch(int64_t count, const int32_t* n, const int32_t* nrhs, const T* const* LU,
                          const int64_t* const* piv, const T* const* X, const T* const* G,
                          T* const* gB, T* const* gA, int accumulate, hipStream_t s)
{
    if (count < 0 || count > INT_MAX) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    if (!n || !nrhs || !LU || !piv || !X || !G || !gB || !gA) return TMVX_ERR_ARG;
    const unsigned grid = (unsigned)std::min<int64_t>(count, 4096);
    hipLaunchKernelGGL(adjoint_batched_kernel<T>, dim3(grid), dim3(64), 0, s, (int)count, n, nrhs,
                       LU, (const i64* const*)piv, G, gB);
    TMVX_TRY(hip_status(hipGetLastError()));
    const unsigned grid2 = (unsigned)std::min<int64_t>(count * (WS_NMAX / ADJ_KC), 16384);
    hipLaunchKernelGGL(adjoint_grad_kernel<T>, dim3(grid2), dim3(256), 0, s, (int)count, n, nrhs,
                       X, (const T* const*)gB, gA, accumulate);
    return hip_status(hipGetLastError());
}

// Factor every A[i] in place: systems with n <= TMVX_BATCH_NMAX through one
// batched launch, larger ones through the blocked LU, one after another.
// Medium systems (above the register-resident batched kernel's limit, up to
// WG_NMAX): one 1024-thread workgroup factors one system in place, column
// by column, in global memory (L2-resident at these sizes): TMV's pivot
// rule (largest |a|, first row on ties), its underflow rule (the column is
// zeroed from the diagonal down and interchanged with itself), the column
// scaled by the pivot's reciprocal, then the rank-1 update of the trailing
// matrix spread over all threads.  info[sys] = 1-based first zero pivot.
constexpr int WG_NMAX = 1024;

template <typename T>
__global__ void __launch_bounds__(1024)
lu_wg_kernel(const int* __restrict__ order, const int32_t* __restrict__ nv, T* const* A,
             int64_t* const* piv, int32_t* __restrict__ info)
{
    constexpr int NW = 1024 / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    __shared__ T s_v[NW];
    __shared__ int s_i[NW];
    __shared__ int s_p, s_zero;
    __shared__ T s_inv;
    const int sys = order[blockIdx.x];
    const int n = nv[sys];
    T* a = A[sys];
    int64_t* pv = piv[sys];
    int first_zero = 0;
    for (int j = 0; j < n; ++j) {
        T bv = T(-1);
        int bi = INT_MAX;
        for (int i = j + t; i < n; i += 1024) {
            const T v = fabs(a[i + (i64)j * n]);
            if (v > bv) { bv = v; bi = i; }
        }
        wave_argmax(bv, bi);
        if (lane == 0) { s_v[wave] = bv; s_i[wave] = bi; }
        __syncthreads();
        if (t == 0) {
            T v = s_v[0];
            int ii = s_i[0];
            for (int q = 1; q < NW; ++q)
                if (better(s_v[q], s_i[q], v, ii)) { v = s_v[q]; ii = s_i[q]; }
            const bool zero = !(v >= tiny<T>());
            s_zero = zero ? 1 : 0;
            s_p = zero ? j : ii;
            pv[j] = zero ? j : ii;
            if (zero && first_zero == 0) first_zero = j + 1;
        }
        __syncthreads();
        if (s_zero) {
            for (int i = j + t; i < n; i += 1024) a[i + (i64)j * n] = T(0);
            __syncthreads();
            continue;
        }
        const int p = s_p;
        if (p != j)
            for (int c = t; c < n; c += 1024) {
                const T x = a[j + (i64)c * n];
                a[j + (i64)c * n] = a[p + (i64)c * n];
                a[p + (i64)c * n] = x;
            }
        __syncthreads();
        if (t == 0) s_inv = T(1) / a[j + (i64)j * n];
        __syncthreads();
        const T inv = s_inv;
        for (int i = j + 1 + t; i < n; i += 1024) a[i + (i64)j * n] *= inv;
        __syncthreads();
        const int m = n - j - 1;
        const i64 tot = (i64)m * m;
        for (i64 e = t; e < tot; e += 1024) {
            const int i = j + 1 + (int)(e % m), c = j + 1 + (int)(e / m);
            a[i + (i64)c * n] -= a[i + (i64)j * n] * a[j + (i64)c * n];
        }
        __syncthreads();
    }
    if (t == 0) info[sys] = first_zero;
}

template <typename T>
static int getrf_batched_impl(int64_t count, const int32_t* n_host, T* const* A, int64_t* const* piv,
                              int32_t* info, hipStream_t s)
{
    if (count < 0 || count > INT_MAX || (count > 0 && (!n_host || !A || !piv || !info)))
        return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    for (int64_t i = 0; i < count; ++i)
        if (n_host[i] < 0) return TMVX_ERR_ARG;
    std::vector<T*> hA((size_t)count);
    std::vector<int64_t*> hP((size_t)count);
    TMVX_HIP(hipMemcpyAsync(hA.data(), A, sizeof(T*) * count, hipMemcpyDeviceToHost, s));
    TMVX_HIP(hipMemcpyAsync(hP.data(), piv, sizeof(int64_t*) * count, hipMemcpyDeviceToHost, s));
    TMVX_HIP(hipStreamSynchronize(s));
    std::vector<int> small, medium, big;
    for (int64_t i = 0; i < count; ++i)
        (n_host[i] <= TMVX_BATCH_NMAX ? small : n_host[i] <= WG_NMAX ? medium : big)
            .push_back((int)i);
    int rc = TMVX_OK;
    if (!medium.empty()) {
        int* dord = nullptr;
        int32_t* dn = nullptr;
        if (hipMalloc((void**)&dord, sizeof(int) * medium.size()) != hipSuccess ||
            hipMalloc((void**)&dn, sizeof(int32_t) * (size_t)count) != hipSuccess)
            rc = TMVX_ERR_ALLOC;
        if (rc == TMVX_OK &&
            (hipMemcpyAsync(dord, medium.data(), sizeof(int) * medium.size(),
                            hipMemcpyHostToDevice, s) != hipSuccess ||
             hipMemcpyAsync(dn, n_host, sizeof(int32_t) * (size_t)count, hipMemcpyHostToDevice,
                            s) != hipSuccess))
            rc = TMVX_ERR_HIP;
        if (rc == TMVX_OK) {
            hipLaunchKernelGGL(lu_wg_kernel<T>, dim3((unsigned)medium.size()), dim3(1024), 0, s,
                               (const int*)dord, (const int32_t*)dn, A, piv, info);
            rc = hip_status(hipGetLastError());
        }
        if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
        hipFree(dord);
        hipFree(dn);
        if (rc != TMVX_OK) return rc;
    }
    if (!small.empty()) {
        const size_t ns = small.size();
        std::vector<int32_t> sn(ns), snr(ns, 0);
        std::vector<void*> ptrs(2 * ns);
        for (size_t k = 0; k < ns; ++k) {
            sn[k] = n_host[small[k]];
            ptrs[k] = hA[small[k]];
            ptrs[ns + k] = hP[small[k]];
        }
        void** dptr = nullptr;
        int32_t* dinfo = nullptr;     // the sub-batch's info, then its indices
        tmvx_batch_plan plan = nullptr;
        if (hipMalloc((void**)&dptr, sizeof(void*) * 2 * ns) != hipSuccess ||
            hipMalloc((void**)&dinfo, sizeof(int32_t) * 2 * ns) != hipSuccess)
            rc = TMVX_ERR_ALLOC;
        if (rc == TMVX_OK &&
            (hipMemcpyAsync(dptr, ptrs.data(), sizeof(void*) * 2 * ns, hipMemcpyHostToDevice, s) !=
                 hipSuccess ||
             hipMemcpyAsync(dinfo + ns, small.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice,
                            s) != hipSuccess))
            rc = TMVX_ERR_HIP;
        if (rc == TMVX_OK)
            rc = tmvx_batch_plan_create(sizeof(T) == 8 ? TMVX_BATCH_F64 : TMVX_BATCH_F32,
                                        (int64_t)ns, sn.data(), snr.data(), 0, &plan, s);
        if (rc == TMVX_OK) {
            if constexpr (sizeof(T) == 8)
                rc = tmvx_gesv_batched_f64(plan, (double* const*)dptr,
                                           (int64_t* const*)(dptr + ns), nullptr, dinfo, s);
            else
                rc = tmvx_gesv_batched_f32(plan, (float* const*)dptr,
                                           (int64_t* const*)(dptr + ns), nullptr, dinfo, s);
        }
        if (rc == TMVX_OK) {
            hipLaunchKernelGGL(scatter_info_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256),
                               0, s, (int)ns, (const int*)(dinfo + ns), (const int32_t*)dinfo,
                               info);
            rc = hip_status(hipGetLastError());
        }
        if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
        if (plan) tmvx_batch_plan_destroy(plan);
        hipFree(dptr);
        hipFree(dinfo);
    }
    // Systems above WG_NMAX: the blocked LU, one system after another on the
    // caller's stream (each call synchronises to read its info).  An
    // earlier version overlapped them from a pool of host threads, each
    // creating streams, events and stream-ordered allocations; that pool
    // crashed the process once under a profiler and was removed: the
    // blocked LU of one such system already spreads over the whole GPU.
    if (rc == TMVX_OK && !big.empty()) {
        std::vector<int32_t> hinfo(big.size(), 0);
        for (size_t k = 0; k < big.size() && rc == TMVX_OK; ++k) {
            const int i = big[k];
            i64 inf = 0;
            rc = getrf_impl<T>(n_host[i], hA[i], n_host[i], (i64*)hP[i], nullptr, &inf, s);
            hinfo[k] = (int32_t)inf;
        }
        for (size_t k = 0; k < big.size() && rc == TMVX_OK; ++k)
            if (hipMemcpyAsync(info + big[k], &hinfo[k], sizeof(int32_t), hipMemcpyHostToDevice,
                               s) != hipSuccess)
                rc = TMVX_ERR_HIP;
        if (rc == TMVX_OK && hipStreamSynchronize(s) != hipSuccess) rc = TMVX_ERR_HIP;
    }
    return rc;
}

template <typename T>
static int adjoint_recompute_impl(int64_t count, const int32_t* n_host, const int32_t* nrhs_host,
                                  const T* const* A, const T* const* X, const T* const* G,
                                  T* const* gB, T* const* gA, int accumulate, hipStream_t s)
{
    if (count < 0 || count > INT_MAX || (count > 0 && (!n_host || !nrhs_host))) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    std::vector<i64> offA((size_t)count), offP((size_t)count);
    i64 ta = 0, tp = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (n_host[i] < 0 || n_host[i] > WS_NMAX || nrhs_host[i] < 0) return TMVX_ERR_UNSUPPORTED;
        offA[i] = ta; ta += (i64)n_host[i] * n_host[i];
        offP[i] = tp; tp += n_host[i];
    }
    char* ws = nullptr;
    const size_t szA = sizeof(T) * (size_t)std::max<i64>(ta, 1);
    const size_t szP = sizeof(i64) * (size_t)std::max<i64>(tp, 1);
    const size_t szPtr = sizeof(void*) * (size_t)count;
    const size_t szI = sizeof(int32_t) * (size_t)count * 3;
    if (hipMalloc((void**)&ws, szA + szP + 2 * szPtr + szI) != hipSuccess) return TMVX_ERR_ALLOC;
    T* LU = reinterpret_cast<T*>(ws);
    i64* P = reinterpret_cast<i64*>(ws + szA);
    void** dLU = reinterpret_cast<void**>(ws + szA + szP);
    void** dP = reinterpret_cast<void**>(ws + szA + szP + szPtr);
    int32_t* dn = reinterpret_cast<int32_t*>(ws + szA + szP + 2 * szPtr);
    int32_t* dnr = dn + count;
    int32_t* dinfo = dn + 2 * count;
    std::vector<void*> hp(2 * (size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        hp[i] = LU + offA[i];
        hp[count + i] = P + offP[i];
    }
    int rc = TMVX_OK;
    if (hipMemcpyAsync(dLU, hp.data(), 2 * szPtr, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dn, n_host, sizeof(int32_t) * count, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(dnr, nrhs_host, sizeof(int32_t) * count, hipMemcpyHostToDevice, s) !=
            hipSuccess)
        rc = TMVX_ERR_HIP;
    if (rc == TMVX_OK) {
        hipLaunchKernelGGL(batched_copy_kernel, dim3((unsigned)std::min<int64_t>(count, 4096)),
                           dim3(256), 0, s, (int)count, (const int*)dn, (const char* const*)A,
                           (char* const*)dLU, (int)sizeof(T));
        rc = hip_status(hipGetLastError());
    }
    if (rc == TMVX_OK)
        rc = getrf_batched_impl<T>(count, n_host, (T* const*)dLU, (int64_t* const*)dP, dinfo, s);
    if (rc == TMVX_OK)
        rc = adjoint_launch<T>(count, dn, dnr, (const T* const*)dLU, (const int64_t* const*)dP, X, G,
                               gB, gA, accumulate, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == TMVX_OK) rc = TMVX_ERR_HIP;
    hipFree(ws);
    return rc;
}

template <typename T>
static int adjoint_single(int64_t n, int64_t nrhs, const T* LU, int64_t lda, const int64_t* piv,
                          const T* X, int64_t ldx, const T* G, int64_t ldg, T* gB, int64_t ldgb,
                          T* gA, int64_t ldga, int accumulate, hipStream_t s)
{
    if (n < 0 || nrhs < 0) return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!LU || !piv || !X || !G || !gB || !gA || lda < n || ldx < n || ldg < n || ldgb < n ||
        ldga < n)
        return TMVX_ERR_ARG;
    if (G != gB)
        TMVX_HIP(hipMemcpy2DAsync(gB, sizeof(T) * ldgb, G, sizeof(T) * ldg, sizeof(T) * n, nrhs,
                                  hipMemcpyDeviceToDevice, s));
    // check = 1: a singular factor or an expired bounded wait in the
    // transposed solve is reported, not turned into a silent gradient.
    int rc;
    if constexpr (sizeof(T) == 8)
        rc = tmvx_getrs_f64(1, n, nrhs, LU, lda, piv, gB, ldgb, 1, s);
    else
        rc = tmvx_getrs_f32(1, n, nrhs, LU, lda, piv, gB, ldgb, 1, s);
    if (rc != TMVX_OK) return rc;
    // gradA (+)= -gB X^T: X (n x nrhs, column-major) read as the row-major
    // nrhs x n operand X^T
    return gemm_launch<T>(0, 1, n, n, nrhs, T(-1), gB, ldgb, X, ldx, accumulate ? T(1) : T(0),
                          gA, ldga, s);
}

} // namespace tmvx

using namespace tmvx;

extern "C" {

int tmvx_gesv_adjoint_f64(int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                          const int64_t* piv, const double* X, int64_t ldx, const double* G,
                          int64_t ldg, double* gradB, int64_t ldgb, double* gradA, int64_t ldga,
                          int accumulate, void* stream)
{
    return adjoint_single<double>(n, nrhs, LU, lda, piv, X, ldx, G, ldg, gradB, ldgb, gradA, ldga,
                                  accumulate, (hipStream_t)stream);
}

int tmvx_gesv_adjoint_f32(int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                          const int64_t* piv, const float* X, int64_t ldx, const float* G,
                          int64_t ldg, float* gradB, int64_t ldgb, float* gradA, int64_t ldga,
                          int accumulate, void* stream)
{
    return adjoint_single<float>(n, nrhs, LU, lda, piv, X, ldx, G, ldg, gradB, ldgb, gradA, ldga,
                                 accumulate, (hipStream_t)stream);
}

int tmvx_gesv_adjoint_batched_f64(int64_t count, const int32_t* n, const int32_t* nrhs,
                                  const double* const* LU, const int64_t* const* piv,
                                  const double* const* X, const double* const* G,
                                  double* const* gradB, double* const* gradA, int accumulate,
                                  void* stream)
{
    return adjoint_launch<double>(count, n, nrhs, LU, piv, X, G, gradB, gradA, accumulate,
                                  (hipStream_t)stream);
}

int tmvx_gesv_adjoint_batched_f32(int64_t count, const int32_t* n, const int32_t* nrhs,
                                  const float* const* LU, const int64_t* const* piv,
                                  const float* const* X, const float* const* G,
                                  float* const* gradB, float* const* gradA, int accumulate,
                                  void* stream)
{
    return adjoint_launch<float>(count, n, nrhs, LU, piv, X, G, gradB, gradA, accumulate,
                                 (hipStream_t)stream);
}

int tmvx_gesv_adjoint_batched_recompute_f64(int64_t count, const int32_t* n_host,
                                            const int32_t* nrhs_host, const double* const* A,
                                            const double* const* X, const double* const* G,
                                            double* const* gradB, double* const* gradA,
                                            int accumulate, void* stream)
{
    return adjoint_recompute_impl<double>(count, n_host, nrhs_host, A, X, G, gradB, gradA,
                                          accumulate, (hipStream_t)stream);
}

int tmvx_gesv_adjoint_batched_recompute_f32(int64_t count, const int32_t* n_host,
                                            const int32_t* nrhs_host, const float* const* A,
                                            const float* const* X, const float* const* G,
                                            float* const* gradB, float* const* gradA,
                                            int accumulate, void* stream)
{
    return adjoint_recompute_impl<float>(count, n_host, nrhs_host, A, X, G, gradB, gradA,
                                         accumulate, (hipStream_t)stream);
}

int tmvx_getrf_batched_f64(int64_t count, const int32_t* n_host, double* const* A,
                           int64_t* const* piv, int32_t* info, void* stream)
{
    return getrf_batched_impl<double>(count, n_host, A, piv, info, (hipStream_t)stream);
}

int tmvx_getrf_batched_f32(int64_t count, const int32_t* n_host, float* const* A,
                           int64_t* const* piv, int32_t* info, void* stream)
{
    return getrf_batched_impl<float>(count, n_host, A, piv, info, (hipStream_t)stream);
}

} // extern "C"
