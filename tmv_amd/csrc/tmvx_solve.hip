// Triangular solves, LU solves and row interchanges for gfx950.
//
// Counterpart of TMV's LU_LDivEq / LU_RDivEq (src/TMV_LUDiv.cpp:185-226,
// 369-410), NonBlasTriLDivEq (src/TMV_TriDiv_M.cpp:190-266; vector column
// sweeps src/TMV_TriDiv_V.cpp:147-218, 321-383) and permuteRows /
// reversePermuteRows (src/TMV_Matrix.cpp:865-930).
//
// op(A) X = B with many right-hand sides is solved by blocked substitution
// over diagonal blocks of TRSM2_NB = 128 rows.  W holds the running
// right-hand side (a copy of B, optionally gathered through a row
// permutation); for diagonal block [b0, b1):
//   - trsm_diag_kernel stages the block of op(A) in LDS and solves it for
//     every column (one wave per column, x_k broadcast by shuffle; like TMV
//     it skips zero entries and flags a zero diagonal met by a non-zero
//     entry), writing X_b to W and to B (optionally scattered through a row
//     permutation);
//   - the rows still to come take the block's contribution on the MFMA
//     GEMM: W[rest, :] -= op(A)[rest, b0:b1] X_b.
#include "tmvx_lu.h"

namespace tmvx {

// Diagonal block of the blocked solve: op(A)[b0:b0+bs, b0:b0+bs] X = W[b0:b0+bs, :]
// for every right-hand side; bs <= TRSM2_NB.  The block is staged in LDS
// once per workgroup; one wave per right-hand side column, lane l holding
// rows l and l + 64, column sweep with the solved entry broadcast by a
// shuffle (zero entries skipped, a zero diagonal met by a non-zero entry
// flagged).  X goes back to W and to B (optionally scattered through a row
// permutation).
constexpr int TRSM2_NB = 128;

template <typename T>
__global__ void __launch_bounds__(256)
trsm_diag_kernel(int lower, int trans, int unit, i64 nrhs, const T* __restrict__ A, i64 lda,
                 T* __restrict__ W, i64 ldw, T* __restrict__ B, i64 ldb,
                 const i64* __restrict__ out_perm, i64 b0, int bs, int* singular)
{
    __shared__ T D[TRSM2_NB][TRSM2_NB + 1];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int e = t; e < TRSM2_NB * TRSM2_NB; e += blockDim.x) {
        const int i = e % TRSM2_NB, j = e / TRSM2_NB;
        T v = T(0);
        if (i < bs && j < bs) v = trans ? A[(b0 + j) + (b0 + i) * lda] : A[(b0 + i) + (b0 + j) * lda];
        D[i][j] = v;
    }
    __syncthreads();
    const i64 col = (i64)blockIdx.x * 4 + wave;
    if (col >= nrhs) return;                 // no barrier below
    T* w = W + col * ldw + b0;
    T x0 = lane < bs ? w[lane] : T(0);
    T x1 = lane + 64 < bs ? w[lane + 64] : T(0);
    for (int q = 0; q < bs; ++q) {
        const int k = lower ? q : bs - 1 - q;
        T xk = __shfl(k < 64 ? x0 : x1, k & 63, 64);
        if (xk != T(0)) {
            if (!unit) {
                const T d = D[k][k];
                if (d == T(0)) {
                    if (lane == 0) atomicOr(singular, 1);
                } else {
                    xk = xk / d;
                }
                if (lane == (k & 63)) { if (k < 64) x0 = xk; else x1 = xk; }
            }
            if (lower) {
                if (lane > k && lane < bs) x0 -= D[lane][k] * xk;
                if (lane + 64 > k && lane + 64 < bs) x1 -= D[lane + 64][k] * xk;
            } else {
                if (lane < k) x0 -= D[lane][k] * xk;
                if (lane + 64 < k) x1 -= D[lane + 64][k] * xk;
            }
        }
    }
    if (lane < bs) {
        w[lane] = x0;
        const i64 orow = out_perm ? out_perm[b0 + lane] : b0 + lane;
        B[orow + col * ldb] = x0;
    }
    if (lane + 64 < bs) {
        w[lane + 64] = x1;
        const i64 orow = out_perm ? out_perm[b0 + lane + 64] : b0 + lane + 64;
        B[orow + col * ldb] = x1;
    }
}

// D[i, j] = S[perm ? perm[i] : i, j]
template <typename T>
__global__ void gather_rows_kernel(i64 m, i64 n, const T* __restrict__ S, i64 lds,
                                   const i64* __restrict__ perm, T* __restrict__ D, i64 ldd)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i < m && j < n) {
        i64 src = perm ? perm[i] : i;
        if (src < 0 || src >= m) src = i;
        D[i + j * ldd] = S[src + j * lds];
    }
}

template <typename T>
int trsm_impl(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A, i64 lda, T* B,
              i64 ldb, const i64* in_perm, const i64* out_perm, int* d_singular, hipStream_t s)
{
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1))
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!A || !B) return TMVX_ERR_ARG;
    if (nrhs > 65535) return TMVX_ERR_UNSUPPORTED;
    const int lower = (upper == 0) != (trans != 0);   // op(A) lower?
    T* W = nullptr;
    if (hipMallocAsync((void**)&W, sizeof(T) * (size_t)n * (size_t)nrhs, s) != hipSuccess)
        return TMVX_ERR_ALLOC;
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3((unsigned)((n + 255) / 256), (unsigned)nrhs),
                       dim3(256), 0, s, n, nrhs, B, ldb, in_perm, W, n);
    int rc = hip_status(hipGetLastError());
    // Blocks of TRSM2_NB rows: the diagonal block by trsm_diag_kernel, then
    // the rows still to come take the block's contribution in one MFMA GEMM,
    // W[rest] -= op(A)[rest, blk] X_blk.
    constexpr int NB = TRSM2_NB;
    const i64 nblk = (n + NB - 1) / NB;
    const unsigned gd = (unsigned)((nrhs + 3) / 4);
    for (i64 q = 0; q < nblk && rc == TMVX_OK; ++q) {
        const i64 blk = lower ? q : nblk - 1 - q;
        const i64 b0 = blk * NB;
        const int bs = (int)((n - b0) < NB ? (n - b0) : NB);
        hipLaunchKernelGGL((trsm_diag_kernel<T>), dim3(gd), dim3(256), 0, s, lower, trans, unit,
                           nrhs, A, lda, W, n, B, ldb, out_perm, b0, bs, d_singular);
        if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
        const i64 r0 = lower ? b0 + bs : 0;          // rows still to come: [r0, r1)
        const i64 r1 = lower ? n : b0;
        if (r1 <= r0) continue;
        // op(A)[r0:r1, b0:b0+bs]: column-major block of A, or (trans) the
        // transpose of A[b0:b0+bs, r0:r1] read row-major
        const T* Ab = trans ? A + b0 + r0 * lda : A + r0 + b0 * lda;
        rc = gemm_launch<T>(trans ? 1 : 0, 0, r1 - r0, nrhs, bs, T(-1), Ab, lda, W + b0, n, T(1),
                            W + r0, n, s);
    }
    hipFreeAsync(W, s);
    return rc;
}

// Right-hand sides up to this count use the single-launch solve (a build
// option for timing the crossover: -DTMVX_TRSV_MAX_RHS=k).
#ifndef TMVX_TRSV_MAX_RHS
#define TMVX_TRSV_MAX_RHS 16
#endif
constexpr i64 TRSV_MAX_RHS = TMVX_TRSV_MAX_RHS;

template <typename T>
int getrs_perm(int trans, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* perm, T* B,
               i64 ldb, int* d_singular, hipStream_t s, int inv)
{
    // A = P^T L U with (P A)[i] = A[perm[i]].
    if (nrhs <= TRSV_MAX_RHS) {
        T* W = nullptr;
        if (hipMallocAsync((void**)&W, sizeof(T) * (size_t)n * (size_t)nrhs, s) != hipSuccess)
            return TMVX_ERR_ALLOC;
        int rc;
        if (!trans) {
            rc = trsv_multi<T>(0, 0, 1, n, nrhs, LU, lda, B, ldb, perm, W, n, nullptr,
                               d_singular, s, inv);
            if (rc == TMVX_OK)
                rc = trsv_multi<T>(1, 0, 0, n, nrhs, LU, lda, W, n, nullptr, B, ldb, nullptr,
                                   d_singular, s, inv);
        } else {
            rc = trsv_multi<T>(1, 1, 0, n, nrhs, LU, lda, B, ldb, nullptr, W, n, nullptr,
                               d_singular, s, inv);
            if (rc == TMVX_OK)
                rc = trsv_multi<T>(0, 1, 1, n, nrhs, LU, lda, W, n, nullptr, B, ldb, perm,
                                   d_singular, s, inv);
        }
        hipFreeAsync(W, s);
        return rc;
    }
    if (!trans) {
        TMVX_TRY(trsm_impl<T>(0, 0, 1, n, nrhs, LU, lda, B, ldb, perm, nullptr, d_singular, s));
        return trsm_impl<T>(1, 0, 0, n, nrhs, LU, lda, B, ldb, nullptr, nullptr, d_singular, s);
    }
    TMVX_TRY(trsm_impl<T>(1, 1, 0, n, nrhs, LU, lda, B, ldb, nullptr, nullptr, d_singular, s));
    return trsm_impl<T>(0, 1, 1, n, nrhs, LU, lda, B, ldb, nullptr, perm, d_singular, s);
}

template <typename T>
int getrs_impl(int trans, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* piv, T* B,
               i64 ldb, int check, hipStream_t s)
{
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1))
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!LU || !piv || !B) return TMVX_ERR_ARG;
    char* ws = nullptr;
    const size_t szperm = sizeof(i64) * (size_t)n;
    if (hipMallocAsync((void**)&ws, szperm + 64, s) != hipSuccess) return TMVX_ERR_ALLOC;
    i64* perm = reinterpret_cast<i64*>(ws);
    int* d_sing = reinterpret_cast<int*>(ws + szperm);
    int rc = zero_bytes(d_sing, sizeof(int), s);
    if (rc == TMVX_OK) rc = piv_to_perm(n, piv, perm, s);
    if (rc == TMVX_OK) rc = getrs_perm<T>(trans, n, nrhs, LU, lda, perm, B, ldb, d_sing, s);
    int hs = 0;
    if (check && rc == TMVX_OK &&
        hipMemcpyAsync(&hs, d_sing, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(ws, s);
    if (check) {
        if (hipStreamSynchronize(s) != hipSuccess) return TMVX_ERR_HIP;
        if (rc == TMVX_OK && (hs & 2)) return TMVX_ERR_TIMEOUT;
        if (rc == TMVX_OK && hs) return TMVX_ERR_SINGULAR;
    }
    return rc;
}

// Interchanges piv[k1..k2) applied in order (forward) or in reverse to the
// rows of each of ncols columns (one thread per column).  Entries outside
// [0, nrows) are skipped.
template <typename T>
__global__ void laswp_seq_kernel(T* __restrict__ A, i64 lda, i64 ncols, i64 k1, i64 k2,
                                 const i64* __restrict__ piv, int forward, i64 nrows)
{
    const i64 col = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncols) return;
    T* a = A + col * lda;
    if (forward) {
        for (i64 k = k1; k < k2; ++k) {
            const i64 p = piv[k];
            if (p != k && p >= 0 && p < nrows) { T x = a[k]; a[k] = a[p]; a[p] = x; }
        }
    } else {
        for (i64 k = k2 - 1; k >= k1; --k) {
            const i64 p = piv[k];
            if (p != k && p >= 0 && p < nrows) { T x = a[k]; a[k] = a[p]; a[p] = x; }
        }
    }
}

template <typename T>
int laswp_impl(int forward, i64 ncols, T* A, i64 lda, i64 k1, i64 k2, const i64* piv,
               hipStream_t s)
{
    if (ncols < 0 || k1 < 0 || k2 < k1 || lda < 1 || k2 > lda) return TMVX_ERR_ARG;
    if (ncols == 0 || k1 == k2) return TMVX_OK;
    if (!A || !piv) return TMVX_ERR_ARG;
    hipLaunchKernelGGL(laswp_seq_kernel<T>, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0,
                       s, A, lda, ncols, k1, k2, piv, forward, lda);
    return hip_status(hipGetLastError());
}

// The public triangular solve: synchronises to report a zero diagonal met
// by a non-zero entry as TMVX_ERR_SINGULAR (TMV throws SingularUpperTriMatrix
// / SingularLowerTriMatrix, src/TMV_TriDiv_V.cpp:171-217).  Few right-hand
// sides take the single-launch solve, as getrs does (n = 16384, 1 RHS:
// ~1 ms against ~10 ms for the blocked form's 128 launch pairs).
template <typename T>
int trsm_public(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A, i64 lda, T* B,
                i64 ldb, hipStream_t s)
{
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1))
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!A || !B) return TMVX_ERR_ARG;
    int* d = nullptr;
    if (hipMallocAsync((void**)&d, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    int rc = zero_bytes(d, sizeof(int), s);
    if (rc == TMVX_OK && nrhs <= TRSV_MAX_RHS)
        rc = trsv_multi<T>(upper, trans, unit, n, nrhs, A, lda, B, ldb, nullptr, B, ldb, nullptr,
                           d, s);
    else if (rc == TMVX_OK)
        rc = trsm_impl<T>(upper, trans, unit, n, nrhs, A, lda, B, ldb, nullptr, nullptr, d, s);
    int hs = 0;
    if (rc == TMVX_OK &&
        hipMemcpyAsync(&hs, d, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(d, s);
    if (hipStreamSynchronize(s) != hipSuccess) return TMVX_ERR_HIP;
    if (rc == TMVX_OK && (hs & 2)) return TMVX_ERR_TIMEOUT;
    if (rc == TMVX_OK && hs) return TMVX_ERR_SINGULAR;
    return rc;
}

// The mixed solver's correction solve (inverted 64 x 64 diagonal blocks,
// tmvx_trsv.hip trsv_inv_kernel) on a whole triangle, for tests and
// experiments: synchronises; a zero diagonal is TMVX_ERR_SINGULAR.
template <typename T>
int trsm_blockinv(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A, i64 lda, T* B,
                  i64 ldb, hipStream_t s)
{
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1))
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!A || !B) return TMVX_ERR_ARG;
    int* d = nullptr;
    if (hipMallocAsync((void**)&d, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    int rc = zero_bytes(d, sizeof(int), s);
    if (rc == TMVX_OK)
        rc = trsv_multi<T>(upper, trans, unit, n, nrhs, A, lda, B, ldb, nullptr, B, ldb, nullptr,
                           d, s, 1);
    int hs = 0;
    if (rc == TMVX_OK &&
        hipMemcpyAsync(&hs, d, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(d, s);
    if (hipStreamSynchronize(s) != hipSuccess) return TMVX_ERR_HIP;
    if (rc == TMVX_OK && (hs & 2)) return TMVX_ERR_TIMEOUT;
    if (rc == TMVX_OK && hs) return TMVX_ERR_SINGULAR;
    return rc;
}

template int trsm_impl<double>(int, int, int, i64, i64, const double*, i64, double*, i64,
                               const i64*, const i64*, int*, hipStream_t);
template int trsm_impl<float>(int, int, int, i64, i64, const float*, i64, float*, i64,
                              const i64*, const i64*, int*, hipStream_t);
template int getrs_perm<double>(int, i64, i64, const double*, i64, const i64*, double*, i64,
                                int*, hipStream_t, int);
template int getrs_perm<float>(int, i64, i64, const float*, i64, const i64*, float*, i64,
                               int*, hipStream_t, int);

} // namespace tmvx

using tmvx::i64;

extern "C" {

int tmvx_getrs_f64(int trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                   const int64_t* piv, double* B, int64_t ldb, int check, void* stream)
{ return tmvx::getrs_impl<double>(trans, n, nrhs, LU, lda, piv, B, ldb, check, (hipStream_t)stream); }
int tmvx_getrs_f32(int trans, int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                   const int64_t* piv, float* B, int64_t ldb, int check, void* stream)
{ return tmvx::getrs_impl<float>(trans, n, nrhs, LU, lda, piv, B, ldb, check, (hipStream_t)stream); }

int tmvx_trsm_f64(int upper, int trans, int unit, int64_t n, int64_t nrhs, const double* A,
                  int64_t lda, double* B, int64_t ldb, void* stream)
{ return tmvx::trsm_public<double>(upper, trans, unit, n, nrhs, A, lda, B, ldb, (hipStream_t)stream); }
int tmvx_trsm_f32(int upper, int trans, int unit, int64_t n, int64_t nrhs, const float* A,
                  int64_t lda, float* B, int64_t ldb, void* stream)
{ return tmvx::trsm_public<float>(upper, trans, unit, n, nrhs, A, lda, B, ldb, (hipStream_t)stream); }

int tmvx_trsm_blockinv_f64(int upper, int trans, int unit, int64_t n, int64_t nrhs,
                           const double* A, int64_t lda, double* B, int64_t ldb, void* stream)
{ return tmvx::trsm_blockinv<double>(upper, trans, unit, n, nrhs, A, lda, B, ldb, (hipStream_t)stream); }
int tmvx_trsm_blockinv_f32(int upper, int trans, int unit, int64_t n, int64_t nrhs,
                           const float* A, int64_t lda, float* B, int64_t ldb, void* stream)
{ return tmvx::trsm_blockinv<float>(upper, trans, unit, n, nrhs, A, lda, B, ldb, (hipStream_t)stream); }

int tmvx_laswp_f64(int forward, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                   const int64_t* piv, void* stream)
{ return tmvx::laswp_impl<double>(forward, ncols, A, lda, k1, k2, piv, (hipStream_t)stream); }
int tmvx_laswp_f32(int forward, int64_t ncols, float* A, int64_t lda, int64_t k1, int64_t k2,
                   const int64_t* piv, void* stream)
{ return tmvx::laswp_impl<float>(forward, ncols, A, lda, k1, k2, piv, (hipStream_t)stream); }

} // extern "C"
