// Triangular solves, LU solves and row interchanges for gfx950.
//
// Counterpart of TMV's LU_LDivEq / LU_RDivEq (src/TMV_LUDiv.cpp:185-226,
// 369-410), NonBlasTriLDivEq (src/TMV_TriDiv_M.cpp:190-266; vector column
// sweeps src/TMV_TriDiv_V.cpp:147-218, 321-383) and permuteRows /
// reversePermuteRows (src/TMV_Matrix.cpp:865-930).
//
// op(A) X = B is solved by blocked substitution over diagonal blocks of
// TRSM_NB rows, one launch per block.  W holds the running right-hand side
// (a copy of B, optionally gathered through a row permutation); the step
// kernel for diagonal block [b0, b1):
//   - every workgroup loads the diagonal block of op(A) and W[b0:b1, :]
//     into LDS and solves it by substitution (one wave per RHS column,
//     x_k broadcast by shuffle; like TMV it skips zero entries and flags a
//     zero diagonal met by a non-zero entry);
//   - workgroup 0 writes X_b to B (optionally scattered through a row
//     permutation);
//   - each workgroup updates its slab of the remaining rows,
//     W[i, :] -= op(A)[i, b0:b1] * X_b.
#include "tmvx_lu.h"

namespace tmvx {

constexpr int TRSM_NB = 32;
constexpr int TRSM_ROWS = 256;   // trailing rows per workgroup (one per thread)

template <typename T, int NB>
__global__ void __launch_bounds__(256)
trsm_step_kernel(int lower, int trans, int unit, i64 n, i64 nrhs,
                 const T* __restrict__ A, i64 lda, T* __restrict__ W, i64 ldw,
                 T* __restrict__ B, i64 ldb, const i64* __restrict__ out_perm,
                 i64 b0, int bs, int* singular)
{
    __shared__ T D[NB][NB + 1];
    __shared__ T X[NB][NB + 1];       // X[i][jj] for a chunk of NB RHS columns
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    auto opA = [&](i64 i, i64 j) -> T {
        return trans ? A[j + i * lda] : A[i + j * lda];
    };
    for (int e = t; e < NB * NB; e += blockDim.x) {
        const int i = e % NB, j = e / NB;
        D[i][j] = (i < bs && j < bs) ? opA(b0 + i, b0 + j) : T(0);
    }
    const i64 trail_lo = lower ? b0 + bs : 0;
    const i64 trail_hi = lower ? n : b0;
    const i64 row = trail_lo + (i64)blockIdx.x * TRSM_ROWS + t;
    const bool has_row = row < trail_hi;
    T arow[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k)
        arow[k] = (has_row && k < bs) ? opA(row, b0 + k) : T(0);

    for (i64 j0 = 0; j0 < nrhs; j0 += NB) {
        const int jn = (int)((nrhs - j0) < NB ? (nrhs - j0) : NB);
        __syncthreads();
        for (int e = t; e < NB * NB; e += blockDim.x) {
            const int i = e % NB, jj = e / NB;
            X[i][jj] = (i < bs && jj < jn) ? W[(b0 + i) + (j0 + jj) * ldw] : T(0);
        }
        __syncthreads();
        for (int jj = wave; jj < jn; jj += 4) {
            T x = lane < bs ? X[lane][jj] : T(0);
            if (lower) {
                for (int k = 0; k < bs; ++k) {
                    T xk = __shfl(x, k, 64);
                    if (!unit && xk != T(0)) {
                        const T d = D[k][k];
                        if (d == T(0)) { if (lane == 0) atomicOr(singular, 1); }
                        else xk = xk / d;
                        if (lane == k) x = xk;
                    }
                    if (lane > k && lane < bs) x -= D[lane][k] * xk;
                }
            } else {
                for (int k = bs - 1; k >= 0; --k) {
                    T xk = __shfl(x, k, 64);
                    if (!unit && xk != T(0)) {
                        const T d = D[k][k];
                        if (d == T(0)) { if (lane == 0) atomicOr(singular, 1); }
                        else xk = xk / d;
                        if (lane == k) x = xk;
                    }
                    if (lane < k) x -= D[lane][k] * xk;
                }
            }
            if (lane < bs) X[lane][jj] = x;
        }
        __syncthreads();
        if (blockIdx.x == 0) {
            for (int e = t; e < NB * NB; e += blockDim.x) {
                const int i = e % NB, jj = e / NB;
                if (i < bs && jj < jn) {
                    const i64 orow = out_perm ? out_perm[b0 + i] : b0 + i;
                    B[orow + (j0 + jj) * ldb] = X[i][jj];
                }
            }
        }
        if (has_row) {
            for (int jj = 0; jj < jn; ++jj) {
                T s = W[row + (j0 + jj) * ldw];
#pragma unroll
                for (int k = 0; k < NB; ++k) s -= arow[k] * X[k][jj];
                W[row + (j0 + jj) * ldw] = s;
            }
        }
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
    constexpr int NB = TRSM_NB;
    const i64 nblk = (n + NB - 1) / NB;
    for (i64 q = 0; q < nblk && rc == TMVX_OK; ++q) {
        const i64 blk = lower ? q : nblk - 1 - q;
        const i64 b0 = blk * NB;
        const int bs = (int)((n - b0) < NB ? (n - b0) : NB);
        const i64 ntrail = lower ? n - b0 - bs : b0;
        const unsigned grid = (unsigned)(ntrail > 0 ? (ntrail + TRSM_ROWS - 1) / TRSM_ROWS : 1);
        hipLaunchKernelGGL((trsm_step_kernel<T, NB>), dim3(grid), dim3(256), 0, s, lower,
                           trans, unit, n, nrhs, A, lda, W, n, B, ldb, out_perm, b0, bs,
                           d_singular);
        rc = hip_status(hipGetLastError());
    }
    hipFreeAsync(W, s);
    return rc;
}

// Right-hand sides up to this count use the single-launch solve.
constexpr i64 TRSV_MAX_RHS = 16;

template <typename T>
int getrs_perm(int trans, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* perm, T* B,
               i64 ldb, int* d_singular, hipStream_t s)
{
    // A = P^T L U with (P A)[i] = A[perm[i]].
    if (nrhs <= TRSV_MAX_RHS) {
        T* W = nullptr;
        if (hipMallocAsync((void**)&W, sizeof(T) * (size_t)n * (size_t)nrhs, s) != hipSuccess)
            return TMVX_ERR_ALLOC;
        int rc;
        if (!trans) {
            rc = trsv_multi<T>(0, 0, 1, n, nrhs, LU, lda, B, ldb, perm, W, n, nullptr,
                               d_singular, s);
            if (rc == TMVX_OK)
                rc = trsv_multi<T>(1, 0, 0, n, nrhs, LU, lda, W, n, nullptr, B, ldb, nullptr,
                                   d_singular, s);
        } else {
            rc = trsv_multi<T>(1, 1, 0, n, nrhs, LU, lda, B, ldb, nullptr, W, n, nullptr,
                               d_singular, s);
            if (rc == TMVX_OK)
                rc = trsv_multi<T>(0, 1, 1, n, nrhs, LU, lda, W, n, nullptr, B, ldb, perm,
                                   d_singular, s);
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

template <typename T>
int trsm_public(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A, i64 lda, T* B,
                i64 ldb, hipStream_t s)
{
    int* d = nullptr;
    if (hipMallocAsync((void**)&d, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    int rc = zero_bytes(d, sizeof(int), s);
    if (rc == TMVX_OK)
        rc = trsm_impl<T>(upper, trans, unit, n, nrhs, A, lda, B, ldb, nullptr, nullptr, d, s);
    hipFreeAsync(d, s);
    return rc;
}

template int trsm_impl<double>(int, int, int, i64, i64, const double*, i64, double*, i64,
                               const i64*, const i64*, int*, hipStream_t);
template int trsm_impl<float>(int, int, int, i64, i64, const float*, i64, float*, i64,
                              const i64*, const i64*, int*, hipStream_t);
template int getrs_perm<double>(int, i64, i64, const double*, i64, const i64*, double*, i64,
                                int*, hipStream_t);
template int getrs_perm<float>(int, i64, i64, const float*, i64, const i64*, float*, i64,
                               int*, hipStream_t);

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

int tmvx_laswp_f64(int forward, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                   const int64_t* piv, void* stream)
{ return tmvx::laswp_impl<double>(forward, ncols, A, lda, k1, k2, piv, (hipStream_t)stream); }
int tmvx_laswp_f32(int forward, int64_t ncols, float* A, int64_t lda, int64_t k1, int64_t k2,
                   const int64_t* piv, void* stream)
{ return tmvx::laswp_impl<float>(forward, ncols, A, lda, k1, k2, piv, (hipStream_t)stream); }

} // extern "C"
