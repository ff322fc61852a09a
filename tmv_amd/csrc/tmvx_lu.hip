// Blocked right-looking LU with partial pivoting, triangular solves and
// row interchanges for gfx950.
//
// Counterpart of TMV's LU_Decompose / RecursiveLUDecompose
// (src/TMV_LUDecompose.cpp:217-407, 514-559), LU_LDivEq / LU_RDivEq
// (src/TMV_LUDiv.cpp:185-226, 369-410), NonBlasTriLDivEq (src/TMV_TriDiv_
// M.cpp:190-266), FindMaxAbsElement (src/TMV_Vector.cpp:473-504) and
// permuteRows (src/TMV_Matrix.cpp:865-930).
//
// Per panel of NB columns starting at j0:
//   1. panel_kernel: G workgroups each hold a contiguous slab of the panel's
//      rows in registers (one row per thread).  Per column: local argmax
//      (largest |a|, lowest row on ties, as TMV's strict '>'), candidates
//      published to a slot, ONE grid barrier, every workgroup reduces the G
//      candidates identically, swaps, scales by the reciprocal of the pivot
//      and applies the rank-1 update to its rows, fusing the next column's
//      argmax.  Slots are double-buffered by column parity so one barrier
//      per column suffices.  The underflow rule of TMV (|pivot| < min:
//      zero the column from the diagonal down, interchange with itself)
//      is kept.  Every wait is bounded; an expired wait raises an error
//      flag and the kernel drains.
//   2. laswp_kernel: the panel's interchanges on the columns outside it.
//   3. trsm_u12_kernel: U12 = L11^-1 A12 with L11 in LDS.
//   4. gemm_launch: A22 -= A21 * U12 on MFMA.
#include "tmvx_common.h"
#include <cfloat>
#include <climits>

namespace tmvx {

constexpr int PANEL_NB = 64;        // panel width (columns)
constexpr int PANEL_THREADS = 512;  // rows per workgroup (one per thread)
constexpr unsigned SPIN_LIMIT = 1u << 21;

template <typename T> __device__ __forceinline__ T tiny();
template <> __device__ __forceinline__ double tiny<double>() { return DBL_MIN; }
template <> __device__ __forceinline__ float tiny<float>() { return FLT_MIN; }

// Candidate (|value|, row); better = larger value, then smaller row.
template <typename T>
__device__ __forceinline__ bool better(T v1, int i1, T v0, int i0)
{ return v1 > v0 || (v1 == v0 && i1 < i0); }

template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& i)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        T v2 = __shfl_xor(v, off, 64);
        int i2 = __shfl_xor(i, off, 64);
        if (better(v2, i2, v, i)) { v = v2; i = i2; }
    }
}

// Grid barrier over gridDim.x workgroups: arrive on a monotonically
// increasing counter, wait until it reaches `target`.  Bounded: after
// SPIN_LIMIT polls it sets *err and returns false.
__device__ bool grid_barrier(unsigned* ctr, unsigned target, int* err)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        unsigned spins = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > SPIN_LIMIT ||
                __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

struct PanelWork {
    void* slot_rows;   // [2][G][NB] T
    void* slot_val;    // [2][G] T
    int* slot_idx;     // [2][G]
    void* diag_rows;   // [2][NB] T
    unsigned* ctr;     // barrier counter
    int* err;          // timeout flag
    i64* first_zero;   // 1-based first zero pivot (min), n+1 if none
};

template <typename T, int NB, int NTH>
__global__ void __launch_bounds__(NTH)
panel_kernel(T* __restrict__ A, i64 lda, i64 n, i64 j0, int jb,
             i64* __restrict__ piv, PanelWork wk, unsigned base)
{
    const int G = gridDim.x, w = blockIdx.x, t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    constexpr int NW = NTH / 64;
    const i64 mp = n - j0;
    const int r = w * NTH + t;                 // panel-relative row
    const bool own = r < mp;
    T* slot_rows = static_cast<T*>(wk.slot_rows);
    T* slot_val = static_cast<T*>(wk.slot_val);
    T* diag_rows = static_cast<T*>(wk.diag_rows);

    __shared__ T s_prow[NB], s_orow[NB];
    __shared__ T s_wv[NW];
    __shared__ int s_wi[NW];
    __shared__ int s_p, s_win, s_zero;

    T row[NB];
    T* gp = A + (j0 + r) + j0 * lda;
#pragma unroll
    for (int c = 0; c < NB; ++c) row[c] = (own && c < jb) ? gp[(i64)c * lda] : T(0);

    // Rows are kept rotated: before column k is processed, row[c] holds
    // panel column (k + c) mod NB, so the current column is always row[0]
    // and every register index is static.  Each step rotates by one; the
    // rows (and the published pivot / diagonal rows) share the rotation.
    auto rotate = [&]() {
        const T first = row[0];
#pragma unroll
        for (int c = 0; c < NB - 1; ++c) row[c] = row[c + 1];
        row[NB - 1] = first;
    };

    T cv = own ? fabs(row[0]) : T(-1);    // candidate for column 0
    bool alive = true;
    int nrot = 0;

    for (int k = 0; k < jb && alive; ++k) {
        const int par = k & 1;
        // -- local argmax over rows >= k --------------------------------
        T v = (own && r >= k) ? cv : T(-1);
        int vi = r;
        wave_argmax(v, vi);
        if (lane == 0) { s_wv[wave] = v; s_wi[wave] = vi; }
        __syncthreads();
        if (t == 0) {
            T bv = s_wv[0]; int bi = s_wi[0];
            for (int q = 1; q < NW; ++q)
                if (better(s_wv[q], s_wi[q], bv, bi)) { bv = s_wv[q]; bi = s_wi[q]; }
            slot_val[par * G + w] = bv;
            wk.slot_idx[par * G + w] = bi;
            s_p = bi;
        }
        __syncthreads();
        // -- publish the candidate row and (owner of row k) the diagonal row
        if (own && r == s_p) {
            T* dst = slot_rows + ((i64)par * G + w) * NB;
#pragma unroll
            for (int c = 0; c < NB; ++c) dst[c] = row[c];
        }
        if (own && r == k) {
            T* dst = diag_rows + (i64)par * NB;
#pragma unroll
            for (int c = 0; c < NB; ++c) dst[c] = row[c];
        }
        if (!grid_barrier(wk.ctr, base + (unsigned)(G * (k + 1)), wk.err)) {
            alive = false;
            break;
        }
        // -- global winner (identical in every workgroup) -----------------
        if (wave == 0) {
            T bv = T(-1); int bi = INT_MAX, bw = 0;
            for (int q = lane; q < G; q += 64) {
                T qv = slot_val[par * G + q];
                int qi = wk.slot_idx[par * G + q];
                if (better(qv, qi, bv, bi)) { bv = qv; bi = qi; bw = q; }
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                T v2 = __shfl_xor(bv, off, 64);
                int i2 = __shfl_xor(bi, off, 64);
                int w2 = __shfl_xor(bw, off, 64);
                if (better(v2, i2, bv, bi)) { bv = v2; bi = i2; bw = w2; }
            }
            if (lane == 0) {
                const bool zero = !(bv >= tiny<T>());   // TMV_Underflow
                s_zero = zero ? 1 : 0;
                s_p = zero ? k : bi;
                s_win = bw;
            }
        }
        __syncthreads();
        const int p = s_p;
        if (s_zero) {
            if (own && r >= k) row[0] = T(0);
            if (w == 0 && t == 0) {
                piv[j0 + k] = j0 + k;
                atomicMin(reinterpret_cast<unsigned long long*>(wk.first_zero),
                          (unsigned long long)(j0 + k + 1));
            }
        } else {
            if (t < NB) {
                s_prow[t] = slot_rows[((i64)par * G + s_win) * NB + t];
                if (p != k) s_orow[t] = diag_rows[(i64)par * NB + t];
            }
            __syncthreads();
            if (own && p != k && r == p) {
#pragma unroll
                for (int c = 0; c < NB; ++c) row[c] = s_orow[c];
            }
            if (own && r == k) {
#pragma unroll
                for (int c = 0; c < NB; ++c) row[c] = s_prow[c];
            }
            if (own && r > k) {
                const T inv = T(1) / s_prow[0];
                const T l = row[0] * inv;
                row[0] = l;
                // rotated columns 1 .. NB-1-k are the panel columns k+1 .. NB-1
#pragma unroll
                for (int c = 1; c < NB; ++c)
                    if (c < NB - k) row[c] -= l * s_prow[c];
            }
            if (w == 0 && t == 0) piv[j0 + k] = j0 + p;
        }
        rotate();
        ++nrot;
        cv = own ? fabs(row[0]) : T(-1);
        __syncthreads();
    }
    // Restore column order (jb < NB leaves the rows part-way round).
    for (int q = nrot % NB; q != 0 && q < NB; ++q) rotate();

#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (own && c < jb) gp[(i64)c * lda] = row[c];
}

// Interchanges piv[k1..k2) applied to columns [0, c1) and [c2, ncols) of A
// (all columns when c1 == c2).  forward: k ascending (permuteRows);
// otherwise descending (reversePermuteRows).
template <typename T>
__global__ void laswp_kernel(T* __restrict__ A, i64 lda, i64 c1, i64 c2,
                             i64 ncols, i64 k1, i64 k2,
                             const i64* __restrict__ piv, int forward, i64 nrows)
{
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 nleft = c1, nright = ncols - c2;
    if (idx >= nleft + nright) return;
    const i64 col = idx < nleft ? idx : c2 + (idx - nleft);
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

// U12 = L11^-1 A12, L11 = unit-lower A[j0:j0+jb, j0:j0+jb],
// A12 = A[j0:j0+jb, j0+jb : j0+jb+ncols].  Lane i of a wave owns row i:
// it keeps row i of L11 in registers and, column by column, forward-
// substitutes with x_k broadcast by readlane.  Each wave solves CW columns
// at once (independent chains interleave) and walks `cpw` columns.
constexpr int U12_CW = 4;

template <typename T, int NB>
__global__ void __launch_bounds__(256)
trsm_u12_kernel(T* __restrict__ A, i64 lda, i64 j0, int jb, i64 ncols, int cpw)
{
    static_assert(NB == 64, "one lane per panel row");
    const int lane = threadIdx.x & 63;
    const i64 wave = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const i64 c0 = wave * cpw;
    if (c0 >= ncols) return;
    const i64 c1 = c0 + cpw < ncols ? c0 + cpw : ncols;
    T Lrow[NB];
    const T* arow = A + (j0 + lane) + j0 * lda;
    static_for<0, NB>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        Lrow[k] = (lane < jb && k < lane) ? arow[(i64)k * lda] : T(0);
    });
    T* bcol = A + (j0 + lane) + (j0 + jb) * lda;
    for (i64 c = c0; c < c1; c += U12_CW) {
        T x[U12_CW];
#pragma unroll
        for (int q = 0; q < U12_CW; ++q)
            x[q] = (lane < jb && c + q < c1) ? bcol[(c + q) * lda] : T(0);
        static_for<0, NB - 1>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
#pragma unroll
            for (int q = 0; q < U12_CW; ++q) {
                const T xk = readlane(x[q], k);
                if (lane > k) x[q] -= Lrow[k] * xk;
            }
        });
#pragma unroll
        for (int q = 0; q < U12_CW; ++q)
            if (lane < jb && c + q < c1) bcol[(c + q) * lda] = x[q];
    }
}

__global__ void reset_counter_kernel(unsigned* ctr, int* err, i64* first_zero, i64 n)
{
    *ctr = 0u;
    *err = 0;
    *first_zero = n + 1;
}

template <typename T>
int getrf_impl(i64 n, T* A, i64 lda, i64* piv, i64* info, hipStream_t s)
{
    if (n < 0 || lda < (n > 0 ? n : 1) || !info) return TMVX_ERR_ARG;
    *info = 0;
    if (n == 0) return TMVX_OK;
    if (!A || !piv) return TMVX_ERR_ARG;
    if (n > (i64)INT_MAX / 2) return TMVX_ERR_UNSUPPORTED;
    constexpr int NB = PANEL_NB, NTH = PANEL_THREADS;
    const int Gmax = (int)((n + NTH - 1) / NTH);
    if (Gmax > 256) return TMVX_ERR_UNSUPPORTED;   // > 131072 rows

    // Workspace: slots + counter + flags, one allocation.
    const size_t szrows = sizeof(T) * 2 * (size_t)Gmax * NB;
    const size_t szval = sizeof(T) * 2 * (size_t)Gmax;
    const size_t szidx = sizeof(int) * 2 * (size_t)Gmax;
    const size_t szdiag = sizeof(T) * 2 * NB;
    const size_t total = szrows + szval + szidx + szdiag + 256;
    char* ws = nullptr;
    if (hipMallocAsync((void**)&ws, total, s) != hipSuccess) return TMVX_ERR_ALLOC;
    PanelWork wk;
    wk.slot_rows = ws;
    wk.slot_val = ws + szrows;
    wk.slot_idx = reinterpret_cast<int*>(ws + szrows + szval);
    wk.diag_rows = ws + szrows + szval + szidx;
    char* tail = ws + szrows + szval + szidx + szdiag;
    tail = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(tail) + 63) & ~uintptr_t(63));
    wk.ctr = reinterpret_cast<unsigned*>(tail);
    wk.err = reinterpret_cast<int*>(tail + 64);
    wk.first_zero = reinterpret_cast<i64*>(tail + 128);

    int rc = TMVX_OK;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, wk.ctr, wk.err,
                       wk.first_zero, n);
    unsigned base = 0;
    for (i64 j0 = 0; j0 < n && rc == TMVX_OK; j0 += NB) {
        const int jb = (int)((n - j0) < NB ? (n - j0) : NB);
        const int G = (int)((n - j0 + NTH - 1) / NTH);
        hipLaunchKernelGGL((panel_kernel<T, NB, NTH>), dim3(G), dim3(NTH), 0, s,
                           A, lda, n, j0, jb, piv, wk, base);
        if (hipGetLastError() != hipSuccess) { rc = TMVX_ERR_HIP; break; }
        base += (unsigned)(G * jb);
        const i64 nother = j0 + (n - j0 - jb);
        if (nother > 0) {
            hipLaunchKernelGGL(laswp_kernel<T>, dim3((unsigned)((nother + 255) / 256)),
                               dim3(256), 0, s, A, lda, j0, j0 + jb, n, j0, j0 + jb,
                               piv, 1, n);
            if (hipGetLastError() != hipSuccess) { rc = TMVX_ERR_HIP; break; }
        }
        const i64 nr = n - j0 - jb;
        if (nr > 0) {
            const int cpw = 16;
            const i64 waves = (nr + cpw - 1) / cpw;
            hipLaunchKernelGGL((trsm_u12_kernel<T, NB>), dim3((unsigned)((waves + 3) / 4)),
                               dim3(256), 0, s, A, lda, j0, jb, nr, cpw);
            if (hipGetLastError() != hipSuccess) { rc = TMVX_ERR_HIP; break; }
            rc = gemm_launch<T>(0, 0, nr, nr, jb, T(-1), A + (j0 + jb) + j0 * lda, lda,
                                A + j0 + (j0 + jb) * lda, lda, T(1),
                                A + (j0 + jb) + (j0 + jb) * lda, lda, s);
        }
    }
    int herr = 0;
    i64 fz = n + 1;
    if (hipMemcpyAsync(&herr, wk.err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&fz, wk.first_zero, sizeof(i64), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(ws, s);
    if (hipStreamSynchronize(s) != hipSuccess) return TMVX_ERR_HIP;
    if (rc != TMVX_OK) return rc;
    if (herr) return TMVX_ERR_TIMEOUT;
    *info = fz <= n ? fz : 0;
    return TMVX_OK;
}

// ---------------------------------------------------------------------------
// Triangular solves: blocked substitution, one launch per diagonal block.
// ---------------------------------------------------------------------------
//
// Solves op(A) X = B for op(A) lower (forward) or upper (backward)
// triangular.  W holds the running right-hand side (a copy of B), B gets
// X block by block.  Step kernel for diagonal block [b0, b1):
//   - every workgroup loads the diagonal block of op(A) and W[b0:b1, :] into
//     LDS and solves it by substitution (one wave per RHS column; a zero
//     diagonal met by a non-zero entry sets *singular, as TMV throws);
//   - workgroup 0 writes X_b into B[b0:b1, :];
//   - each workgroup updates its slab of the remaining rows:
//     W[i, :] -= op(A)[i, b0:b1] * X_b.
constexpr int TRSM_NB = 32;
constexpr int TRSM_ROWS = 256;   // trailing rows per workgroup (one per thread)

template <typename T, int NB>
__global__ void __launch_bounds__(256)
trsm_step_kernel(int lower, int trans, int unit, i64 n, i64 nrhs,
                 const T* __restrict__ A, i64 lda, T* __restrict__ W, i64 ldw,
                 T* __restrict__ B, i64 ldb, i64 b0, int bs, int* singular)
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
    // Trailing rows handled by this workgroup.
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
        // Substitution: wave handles columns jj = wave, wave+4, ...
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
                if (i < bs && jj < jn) B[(b0 + i) + (j0 + jj) * ldb] = X[i][jj];
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

template <typename T>
__global__ void copy_kernel(i64 m, i64 n, const T* __restrict__ S, i64 lds,
                            T* __restrict__ D, i64 ldd)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i < m && j < n) D[i + j * ldd] = S[i + j * lds];
}

template <typename T>
int trsm_impl(int upper, int trans, int unit, i64 n, i64 nrhs, const T* A,
              i64 lda, T* B, i64 ldb, int* d_singular, hipStream_t s)
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
    hipLaunchKernelGGL(copy_kernel<T>, dim3((unsigned)((n + 255) / 256), (unsigned)nrhs),
                       dim3(256), 0, s, n, nrhs, B, ldb, W, n);
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
                           trans, unit, n, nrhs, A, lda, W, n, B, ldb, b0, bs, d_singular);
        rc = hip_status(hipGetLastError());
    }
    hipFreeAsync(W, s);
    return rc;
}

template <typename T>
int getrs_impl(int trans, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* piv,
               T* B, i64 ldb, int check, hipStream_t s)
{
    if (n < 0 || nrhs < 0 || lda < (n > 0 ? n : 1) || ldb < (n > 0 ? n : 1))
        return TMVX_ERR_ARG;
    if (n == 0 || nrhs == 0) return TMVX_OK;
    if (!LU || !piv || !B) return TMVX_ERR_ARG;
    int* d_sing = nullptr;
    if (hipMallocAsync((void**)&d_sing, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    if (hipMemsetAsync(d_sing, 0, sizeof(int), s) != hipSuccess) return TMVX_ERR_HIP;
    int rc = TMVX_OK;
    const unsigned gcols = (unsigned)((nrhs + 255) / 256);
    if (!trans) {
        hipLaunchKernelGGL(laswp_kernel<T>, dim3(gcols), dim3(256), 0, s, B, ldb, (i64)0,
                           (i64)0, nrhs, (i64)0, n, piv, 1, n);
        rc = hip_status(hipGetLastError());
        if (rc == TMVX_OK) rc = trsm_impl<T>(0, 0, 1, n, nrhs, LU, lda, B, ldb, d_sing, s);
        if (rc == TMVX_OK) rc = trsm_impl<T>(1, 0, 0, n, nrhs, LU, lda, B, ldb, d_sing, s);
    } else {
        rc = trsm_impl<T>(1, 1, 0, n, nrhs, LU, lda, B, ldb, d_sing, s);
        if (rc == TMVX_OK) rc = trsm_impl<T>(0, 1, 1, n, nrhs, LU, lda, B, ldb, d_sing, s);
        if (rc == TMVX_OK) {
            hipLaunchKernelGGL(laswp_kernel<T>, dim3(gcols), dim3(256), 0, s, B, ldb, (i64)0,
                               (i64)0, nrhs, (i64)0, n, piv, 0, n);
            rc = hip_status(hipGetLastError());
        }
    }
    int hs = 0;
    if (check && rc == TMVX_OK) {
        if (hipMemcpyAsync(&hs, d_sing, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess)
            rc = TMVX_ERR_HIP;
    }
    hipFreeAsync(d_sing, s);
    if (check) {
        if (hipStreamSynchronize(s) != hipSuccess) return TMVX_ERR_HIP;
        if (rc == TMVX_OK && hs) return TMVX_ERR_SINGULAR;
    }
    return rc;
}

template int getrf_impl<double>(i64, double*, i64, i64*, i64*, hipStream_t);
template int getrf_impl<float>(i64, float*, i64, i64*, i64*, hipStream_t);
template int getrs_impl<double>(int, i64, i64, const double*, i64, const i64*, double*, i64, int, hipStream_t);
template int getrs_impl<float>(int, i64, i64, const float*, i64, const i64*, float*, i64, int, hipStream_t);

} // namespace tmvx

using tmvx::i64;

extern "C" {

int tmvx_getrf_f64(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return tmvx::getrf_impl<double>(n, A, lda, piv, info, (hipStream_t)stream); }
int tmvx_getrf_f32(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return tmvx::getrf_impl<float>(n, A, lda, piv, info, (hipStream_t)stream); }

int tmvx_getrs_f64(int trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                   const int64_t* piv, double* B, int64_t ldb, int check, void* stream)
{ return tmvx::getrs_impl<double>(trans, n, nrhs, LU, lda, piv, B, ldb, check, (hipStream_t)stream); }
int tmvx_getrs_f32(int trans, int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                   const int64_t* piv, float* B, int64_t ldb, int check, void* stream)
{ return tmvx::getrs_impl<float>(trans, n, nrhs, LU, lda, piv, B, ldb, check, (hipStream_t)stream); }

int tmvx_trsm_f64(int upper, int trans, int unit, int64_t n, int64_t nrhs, const double* A,
                  int64_t lda, double* B, int64_t ldb, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    int* d = nullptr;
    if (hipMallocAsync((void**)&d, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    if (hipMemsetAsync(d, 0, sizeof(int), s) != hipSuccess) return TMVX_ERR_HIP;
    int rc = tmvx::trsm_impl<double>(upper, trans, unit, n, nrhs, A, lda, B, ldb, d, s);
    hipFreeAsync(d, s);
    return rc;
}
int tmvx_trsm_f32(int upper, int trans, int unit, int64_t n, int64_t nrhs, const float* A,
                  int64_t lda, float* B, int64_t ldb, void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    int* d = nullptr;
    if (hipMallocAsync((void**)&d, sizeof(int), s) != hipSuccess) return TMVX_ERR_ALLOC;
    if (hipMemsetAsync(d, 0, sizeof(int), s) != hipSuccess) return TMVX_ERR_HIP;
    int rc = tmvx::trsm_impl<float>(upper, trans, unit, n, nrhs, A, lda, B, ldb, d, s);
    hipFreeAsync(d, s);
    return rc;
}

int tmvx_laswp_f64(int forward, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                   const int64_t* piv, void* stream)
{
    if (ncols < 0 || k1 < 0 || k2 < k1 || lda < 1) return TMVX_ERR_ARG;
    if (ncols == 0 || k1 == k2) return TMVX_OK;
    hipLaunchKernelGGL(tmvx::laswp_kernel<double>, dim3((unsigned)((ncols + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, A, lda, (i64)0, (i64)0, ncols, k1,
                       k2, piv, forward, lda);
    return tmvx::hip_status(hipGetLastError());
}
int tmvx_laswp_f32(int forward, int64_t ncols, float* A, int64_t lda, int64_t k1, int64_t k2,
                   const int64_t* piv, void* stream)
{
    if (ncols < 0 || k1 < 0 || k2 < k1 || lda < 1) return TMVX_ERR_ARG;
    if (ncols == 0 || k1 == k2) return TMVX_OK;
    hipLaunchKernelGGL(tmvx::laswp_kernel<float>, dim3((unsigned)((ncols + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, A, lda, (i64)0, (i64)0, ncols, k1,
                       k2, piv, forward, lda);
    return tmvx::hip_status(hipGetLastError());
}

} // extern "C"
