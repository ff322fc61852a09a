// Blocked right-looking LU with partial pivoting for gfx950.
//
// Counterpart of TMV's LU_Decompose / RecursiveLUDecompose
// (src/TMV_LUDecompose.cpp:217-407, 514-559), FindMaxAbsElement
// (src/TMV_Vector.cpp:473-504) and permuteRows (src/TMV_Matrix.cpp:865-930).
// TMV recurses on column halves down to 1-2 column base cases; here the
// matrix is processed in panels of NB = 64 columns.  On tie-free inputs both
// choose the same pivots (GEPP's pivot sequence is fixed by the arithmetic
// of exact elimination); factors agree to rounding.
//
// Per panel of NB columns starting at j0:
//   1. panel_kernel: G workgroups each hold a contiguous slab of the panel's
//      rows in registers (one row per thread, kept rotated so the current
//      column is always register 0).  Per column: local argmax (largest
//      |a|, lowest row on ties, as TMV's strict '>'); each workgroup
//      publishes its candidate row, value and index to its slot as tagged
//      64-bit words (32 payload bits + the global column number), and every
//      workgroup polls all slots until each word carries this column's tag
//      -- the grid synchronisation and the data exchange are one round
//      trip.  Every workgroup then reduces the G candidates identically,
//      swaps, scales by the reciprocal of the pivot and updates its rows.
//      Slots are double-buffered by column parity.  TMV's
//      underflow rule (largest |candidate| < min: zero the column from the
//      diagonal down, interchange with itself) is kept.  Every wait is
//      bounded; an expired wait raises an error flag and the kernel drains.
//   2. laswp_panel_kernel: the panel's net row permutation (at most 2*NB
//      rows) computed once per workgroup, applied to every column outside
//      the panel as a one-wave-per-column gather; the composed permutation
//      vector is updated the same way.
//   3. trsm_u12_kernel: U12 = L11^-1 A12 (lane i owns row i of L11).
//   4. gemm_launch: A22 -= A21 * U12 on MFMA, with look-ahead: the next
//      outer panel's columns are updated first on the caller's stream,
//      the rest on a second stream while the next panel is factored.
#include "tmvx_lu.h"
#include "tmvx_pivot.h"
#include <cfloat>
#include <climits>

namespace tmvx {

constexpr int PANEL_NB = 64;        // panel width (columns)
constexpr int PANEL_THREADS = 512;  // rows per workgroup (one per thread)
constexpr int PANEL_GMAX = 64;      // workgroups per panel (n <= 32768)
constexpr unsigned SPIN_LIMIT = 1u << 22;

// Device-coherent (agent scope) relaxed loads/stores for the exchange slots:
// they are performed at the device coherence point, so a workgroup on any
// XCD observes them without cache maintenance.
template <typename T>
__device__ __forceinline__ void cstore(T* p, T v)
{ __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T>
__device__ __forceinline__ T cload(const T* p)
{ return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

struct PanelWork {
    unsigned long long* slots;   // [2][G][SLOT] tagged words (see SlotLayout)
    int* err;                    // timeout flag
    i64* first_zero;             // 1-based first zero pivot (min), n+1 if none
};

// Exchange slot of one workgroup for one column parity, in tagged 64-bit
// words: (column tag << 32) | 32 payload bits.  A double takes two words
// (low, high half), a float one.  Each word is written and read as one
// 64-bit atomic, so a reader that sees the current tag in a word also sees
// that word's payload: no separate flag and no ordering fence is needed.
template <typename T, int NB> struct SlotLayout {
    static constexpr int WPV = sizeof(T) / 4;             // words per value
    static constexpr int ROW = 0;                          // candidate row
    static constexpr int DIAG = NB * WPV;                  // row k (slab 0 only)
    static constexpr int VAL = 2 * NB * WPV;               // candidate |value|
    static constexpr int IDX = VAL + WPV;                  // candidate row index
    static constexpr int SIZE = IDX + 1;
};

template <typename T>
__device__ __forceinline__ unsigned half_bits(T v, int h)
{
    if constexpr (sizeof(T) == 8) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        return h ? (unsigned)(b >> 32) : (unsigned)b;
    } else {
        return (unsigned)__float_as_uint(v);
    }
}

template <typename T>
__device__ __forceinline__ T from_halves(const unsigned* h)
{
    if constexpr (sizeof(T) == 8)
        return __longlong_as_double((long long)(((unsigned long long)h[1] << 32) | h[0]));
    else
        return __uint_as_float(h[0]);
}

// Phase timestamps of the panel loop (shader clock), only in a build with
// -DTMVX_PANEL_TRACE: workgroup 0, thread 0, first traced panel.
#ifdef TMVX_PANEL_TRACE
__device__ unsigned long long* g_panel_trace;
#define PANEL_TRACE(k, ph)                                                     \
    do {                                                                       \
        if (g_panel_trace && w == 0 && t == 0 && j0 == 0)                      \
            g_panel_trace[(k) * 8 + (ph)] = clock64();                         \
    } while (0)
#else
#define PANEL_TRACE(k, ph) do {} while (0)
#endif

template <typename T, int NB, int NTH>
__global__ void __launch_bounds__(NTH)
panel_kernel(T* __restrict__ A, i64 lda, i64 n, i64 j0, int jb,
             i64* __restrict__ piv, PanelWork wk)
{
    typedef SlotLayout<T, NB> SL;
    constexpr int WPV = SL::WPV;
    const int G = gridDim.x, w = blockIdx.x, t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    constexpr int NW = NTH / 64;
    const i64 mp = n - j0;
    const int r = w * NTH + t;                 // panel-relative row
    const bool own = r < mp;

    // Received payload halves: candidate rows of every slab, values,
    // indices, and row k.
    __shared__ unsigned s_cand[PANEL_GMAX][NB * WPV];
    __shared__ unsigned s_diag[NB * WPV];
    __shared__ unsigned s_cval[PANEL_GMAX][WPV];
    __shared__ int s_cidx[PANEL_GMAX];
    __shared__ __attribute__((aligned(16))) T s_out[2][NB];   // rows to publish (cand, diag)
    __shared__ __attribute__((aligned(16))) T s_prow[NB];   // winning row
    __shared__ __attribute__((aligned(16))) T s_orow[NB];   // row k before the swap
    __shared__ T s_wv[NW];
    __shared__ int s_wi[NW];
    __shared__ T s_bv;
    __shared__ int s_p, s_win, s_zero, s_fail;

    T row[NB];
    T* gp = A + (j0 + r) + j0 * lda;
#pragma unroll
    for (int c = 0; c < NB; ++c) row[c] = (own && c < jb) ? gp[(i64)c * lda] : T(0);

    // Rows are kept rotated: before column k is processed, row[c] holds
    // panel column (k + c) mod NB, so the current column is always row[0]
    // and every register index is static.  Each step rotates by one; the
    // published candidate and diagonal rows share the rotation.
    auto rotate = [&]() {
        const T first = row[0];
#pragma unroll
        for (int c = 0; c < NB - 1; ++c) row[c] = row[c + 1];
        row[NB - 1] = first;
    };

    if (t == 0) s_fail = 0;
    T cv = own ? fabs(row[0]) : T(-1);    // candidate for column 0
    bool alive = true;
    int nrot = 0;

    for (int k = 0; k < jb && alive; ++k) {
        const int par = k & 1;
        const unsigned long long tag = (unsigned long long)(j0 + k + 1) << 32;
        unsigned long long* myslot = wk.slots + ((i64)par * G + w) * SL::SIZE;
        PANEL_TRACE(k, 0);
        // -- local argmax over rows >= k --------------------------------
        T v = (own && r >= k) ? cv : T(-1);
        int vi = r;
        wave_argmax(v, vi);
        if (lane == 0) { s_wv[wave] = v; s_wi[wave] = vi; }
        __syncthreads();
        if (wave == 0) {                 // across waves: one more DPP argmax
            T bv = lane < NW ? s_wv[lane] : T(-1);
            int bi = lane < NW ? s_wi[lane] : INT_MAX;
            wave_argmax(bv, bi);
            if (lane == 0) { s_bv = bv; s_p = bi; }
        }
        __syncthreads();
        PANEL_TRACE(k, 1);
        // -- publish the candidate row and (slab 0) row k as tagged words --
        typedef typename Vec16<T>::type VT;
        constexpr int VN = Vec16<T>::n;
        if (own && r == s_p) {          // 16-byte LDS stores
            static_for<0, NB / VN>([&](auto cc) {
                constexpr int c = decltype(cc)::value * VN;
                VT x;
                static_for<0, VN>([&](auto uc) { x[decltype(uc)::value] = row[c + decltype(uc)::value]; });
                *reinterpret_cast<VT*>(&s_out[0][c]) = x;
            });
        }
        if (own && r == k) {
            static_for<0, NB / VN>([&](auto cc) {
                constexpr int c = decltype(cc)::value * VN;
                VT x;
                static_for<0, VN>([&](auto uc) { x[decltype(uc)::value] = row[c + decltype(uc)::value]; });
                *reinterpret_cast<VT*>(&s_out[1][c]) = x;
            });
        }
        __syncthreads();
        for (int e = t; e < SL::SIZE; e += NTH) {
            unsigned pay;
            if (e < SL::DIAG)      pay = half_bits(s_out[0][e / WPV], e % WPV);
            else if (e < SL::VAL) {
                if (w != 0) continue;                       // row k: slab 0 only
                pay = half_bits(s_out[1][(e - SL::DIAG) / WPV], (e - SL::DIAG) % WPV);
            }
            else if (e < SL::IDX)  pay = half_bits(s_bv, e - SL::VAL);
            else                   pay = (unsigned)s_p;
            cstore(myslot + e, tag | pay);
        }
        PANEL_TRACE(k, 2);
        // -- receive every slab's words once they carry this column's tag --
        //    Each thread issues all of its loads, then re-polls only the
        //    words that are still stale: one round trip per poll, not one
        //    per word.
        {
            const int nrow = G * SL::DIAG;               // candidate-row words
            const int total = nrow + SL::DIAG + G * (WPV + 1);
            constexpr int CW = sizeof(T) == 8 ? 4 : 8;   // words per thread per poll
            auto src_of = [&](int e) -> const unsigned long long* {
                if (e < nrow)
                    return wk.slots + ((i64)par * G + e / SL::DIAG) * SL::SIZE + SL::ROW +
                           e % SL::DIAG;
                if (e < nrow + SL::DIAG)
                    return wk.slots + (i64)par * G * SL::SIZE + SL::DIAG + (e - nrow);
                const int f = e - nrow - SL::DIAG, q = f / (WPV + 1), h = f % (WPV + 1);
                return wk.slots + ((i64)par * G + q) * SL::SIZE + SL::VAL + h;
            };
            unsigned spins = 0;
            bool ok = true;
            for (int e0 = t; e0 < total && ok; e0 += CW * NTH) {
                unsigned pend = 0;                        // bit i: word e0 + i*NTH stale
                static_for<0, CW>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    if (e0 + i * NTH < total) pend |= 1u << i;
                });
                while (pend) {
                    unsigned long long wv[CW];
                    static_for<0, CW>([&](auto ic) {
                        constexpr int i = decltype(ic)::value;
                        wv[i] = (pend >> i) & 1u ? cload(src_of(e0 + i * NTH)) : 0ull;
                    });
                    static_for<0, CW>([&](auto ic) {
                        constexpr int i = decltype(ic)::value;
                        if (((pend >> i) & 1u) && (wv[i] & ~0xffffffffull) == tag) {
                            pend &= ~(1u << i);
                            const int e = e0 + i * NTH;
                            const unsigned pay = (unsigned)wv[i];
                            if (e < nrow) s_cand[e / SL::DIAG][e % SL::DIAG] = pay;
                            else if (e < nrow + SL::DIAG) s_diag[e - nrow] = pay;
                            else {
                                const int f = e - nrow - SL::DIAG, q = f / (WPV + 1),
                                          h = f % (WPV + 1);
                                if (h < WPV) s_cval[q][h] = pay; else s_cidx[q] = (int)pay;
                            }
                        }
                    });
                    if (!pend) break;
                    if (++spins > SPIN_LIMIT ||
                        ((spins & 255) == 0 &&
                         __hip_atomic_load(wk.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                        __hip_atomic_store(wk.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        s_fail = 1;
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        PANEL_TRACE(k, 3);
        __syncthreads();
        PANEL_TRACE(k, 4);
        if (s_fail) { alive = false; break; }
        // -- global winner (identical in every workgroup): lane q of wave 0
        //    holds slab q's candidate, one DPP argmax picks the best --------
        if (wave == 0) {
            T bv = T(-1);
            int bi = INT_MAX;
            if (lane < G) { bv = from_halves<T>(s_cval[lane]); bi = s_cidx[lane]; }
            const T myv = bv;
            const int myi = bi;
            wave_argmax(bv, bi);
            // the winning slab: lowest lane holding the winning candidate
            const unsigned long long hit = __ballot(lane < G && myv == bv && myi == bi);
            if (lane == 0) {
                const bool zero = !(bv >= tiny<T>());   // TMV_Underflow
                s_zero = zero ? 1 : 0;
                s_p = zero ? k : bi;
                s_win = hit ? (int)__builtin_ctzll(hit) : 0;
            }
        }
        __syncthreads();
        if (t < NB) s_prow[t] = from_halves<T>(&s_cand[s_win][t * WPV]);
        else if (t < 2 * NB) s_orow[t - NB] = from_halves<T>(&s_diag[(t - NB) * WPV]);
        __syncthreads();
        PANEL_TRACE(k, 5);
        const int p = s_p;
        if (s_zero) {
            if (own && r >= k) row[0] = T(0);
            if (w == 0 && t == 0) {
                piv[j0 + k] = j0 + k;
                atomicMin(reinterpret_cast<unsigned long long*>(wk.first_zero),
                          (unsigned long long)(j0 + k + 1));
            }
        } else {
            if (own && p != k && r == p) {
#pragma unroll
                for (int c = 0; c < NB; ++c) row[c] = s_orow[c];
            }
            if (own && r == k) {
#pragma unroll
                for (int c = 0; c < NB; ++c) row[c] = s_prow[c];
            }
            if (own && r > k) {
                // rotated columns 1 .. NB-1-k are the panel columns k+1 .. NB-1
                if constexpr (sizeof(T) == 4) {
                    T pr[NB];                   // pivot row, 16-byte LDS loads
                    static_for<0, NB / VN>([&](auto cc) {
                        constexpr int c = decltype(cc)::value * VN;
                        const VT x = *reinterpret_cast<const VT*>(&s_prow[c]);
                        static_for<0, VN>([&](auto uc) {
                            pr[c + decltype(uc)::value] = x[decltype(uc)::value];
                        });
                    });
                    const T l = row[0] * (T(1) / pr[0]);
                    row[0] = l;
#pragma unroll
                    for (int c = 1; c < NB; ++c)
                        if (c < NB - k) row[c] -= l * pr[c];
                } else {                        // (registers: the row alone is 128)
                    const T l = row[0] * (T(1) / s_prow[0]);
                    row[0] = l;
#pragma unroll
                    for (int c = 1; c < NB; ++c)
                        if (c < NB - k) row[c] -= l * s_prow[c];
                }
            }
            if (w == 0 && t == 0) piv[j0 + k] = j0 + p;
        }
        PANEL_TRACE(k, 6);
        rotate();
        ++nrot;
        cv = own ? fabs(row[0]) : T(-1);
        __syncthreads();
        PANEL_TRACE(k, 7);
    }
    // Restore column order (jb < NB leaves the rows part-way round).
    for (int q = nrot % NB; q != 0 && q < NB; ++q) rotate();

#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (own && c < jb) gp[(i64)c * lda] = row[c];
}

// Net effect of interchanges piv[j0 .. j0+jb) (absolute rows, each >= its
// step): new row j0+i = old row src[i] (i < jb) and new row odst[q] = old
// row osrc[q] for the nout rows below the block that took part.  Computed
// by one wave: lane i holds src[i] and the i-th outside row, lookups of an
// outside row are a ballot.  Out-of-range pivots are ignored (no access).
template <int NB>
struct PanelPerm {
    int nout;
    i64 src[NB];
    i64 odst[NB], osrc[NB];
};

template <int NB>
__device__ void panel_perm_wave(PanelPerm<NB>& P, const i64* lpiv, i64 j0, int jb, i64 nrows)
{
    static_assert(NB == 64, "one lane per panel column");
    const int lane = threadIdx.x & 63;
    long long src = j0 + lane, odst = -1, osrc = -1;
    int nout = 0;
    for (int k = 0; k < jb; ++k) {
        const i64 p = lpiv[k];
        if (p <= j0 + k || p >= nrows) continue;            // uniform
        if (p < j0 + jb) {
            const int kp = (int)(p - j0);
            const long long a = __shfl(src, k, 64), b = __shfl(src, kp, 64);
            if (lane == k) src = b;
            if (lane == kp) src = a;
        } else {
            const unsigned long long hit = __ballot(lane < nout && odst == p);
            int q;
            if (hit) q = __ffsll((long long)hit) - 1;
            else {
                q = nout++;
                if (lane == q) { odst = p; osrc = p; }
            }
            const long long a = __shfl(src, k, 64), b = __shfl(osrc, q, 64);
            if (lane == k) src = b;
            if (lane == q) osrc = a;
        }
    }
    P.src[lane] = src;
    P.odst[lane] = odst;
    P.osrc[lane] = osrc;
    if (lane == 0) P.nout = nout;
}

// Apply the panel's net permutation to columns [a0, a1) and [b0, b1) of A
// (one wave per column: all loads, then all stores), and to the composed
// permutation vector `perm` (when given) by the first wave of workgroup 0.
template <typename T, int NB>
__global__ void __launch_bounds__(256)
laswp_panel_kernel(T* __restrict__ A, i64 lda, i64 nrows, i64 a0, i64 a1, i64 b0, i64 b1,
                   i64 j0, int jb, const i64* __restrict__ piv, i64* __restrict__ perm,
                   int cols_per_wave)
{
    __shared__ PanelPerm<NB> P;
    __shared__ i64 lpiv[NB];
    if (threadIdx.x < jb) lpiv[threadIdx.x] = piv[j0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < 64) panel_perm_wave<NB>(P, lpiv, j0, jb, nrows);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int nout = P.nout;
    if (perm != nullptr && blockIdx.x == 0 && threadIdx.x < 64) {
        i64 v1 = 0, v2 = 0;
        if (lane < jb) v1 = perm[P.src[lane]];
        if (lane < nout) v2 = perm[P.osrc[lane]];
        __builtin_amdgcn_s_waitcnt(0);     // every load returned before any store
        if (lane < jb) perm[j0 + lane] = v1;
        if (lane < nout) perm[P.odst[lane]] = v2;
    }
    const i64 nleft = a1 - a0, nright = b1 - b0, total = nleft + nright;
    const i64 wave = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int q = 0; q < cols_per_wave; ++q) {
        const i64 idx = wave * cols_per_wave + q;
        if (idx >= total) break;
        const i64 col = idx < nleft ? a0 + idx : b0 + (idx - nleft);
        T* a = A + col * lda;
        T v1 = T(0), v2 = T(0);
        if (lane < jb) v1 = a[P.src[lane]];
        if (lane < nout) v2 = a[P.osrc[lane]];
        __builtin_amdgcn_s_waitcnt(0);     // every load returned before any store
        if (lane < jb) a[j0 + lane] = v1;
        if (lane < nout) a[P.odst[lane]] = v2;
    }
}

// U12 = L11^-1 A12, L11 = unit-lower A[j0:j0+jb, j0:j0+jb],
// A12 = A[j0:j0+jb, c0 : c0+ncols].  Lane i of a wave owns row i:
// it keeps row i of L11 in registers and, column by column, forward-
// substitutes with x_k broadcast by readlane.  Each wave solves CW columns
// at once (independent chains interleave) and walks `cpw` columns.
constexpr int U12_CW = 4;

template <typename T, int NB>
__global__ void __launch_bounds__(256)
trsm_u12_kernel(T* __restrict__ A, i64 lda, i64 j0, int jb, i64 cbeg, i64 ncols, int cpw)
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
    T* bcol = A + (j0 + lane) + cbeg * lda;
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

__global__ void getrf_init_kernel(unsigned long long* tags, i64 ntags, int* err,
                                  i64* first_zero, i64 n, i64* perm)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntags) tags[i] = 0ull;
    if (i == 0) { *err = 0; *first_zero = n + 1; }
    if (perm != nullptr && i < n) perm[i] = i;
}

__global__ void zero_kernel(unsigned char* p, size_t bytes)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < bytes) p[i] = 0;
}

int zero_bytes(void* p, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return TMVX_OK;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((bytes + 255) / 256)), dim3(256), 0, s,
                       static_cast<unsigned char*>(p), bytes);
    return hip_status(hipGetLastError());
}

template <typename T>
int getrf_impl(i64 n, T* A, i64 lda, i64* piv, i64* perm, i64* info, hipStream_t s, int lowprec,
               int tournament)
{
    if (n < 0 || lda < (n > 0 ? n : 1) || !info) return TMVX_ERR_ARG;
    *info = 0;
    if (n == 0) return TMVX_OK;
    if (!A || !piv) return TMVX_ERR_ARG;
    constexpr int NB = PANEL_NB, NTH = PANEL_THREADS;
    const int Gmax = (int)((n + NTH - 1) / NTH);
    if (Gmax > PANEL_GMAX) return TMVX_ERR_UNSUPPORTED;   // n > 32768

    const i64 nslot = 2 * (i64)Gmax * SlotLayout<T, NB>::SIZE;
    const size_t szslots = sizeof(unsigned long long) * (size_t)nslot;
    const bool tslu = sizeof(T) == 4 && tournament != 0;
    const size_t szts = tslu ? sizeof(int) * (size_t)tslu_ws_ints(n) : 0;
    const size_t total = szslots + 256 + szts;
    char* ws = nullptr;
    if (hipMallocAsync((void**)&ws, total, s) != hipSuccess) return TMVX_ERR_ALLOC;
    PanelWork wk;
    wk.slots = reinterpret_cast<unsigned long long*>(ws);
    char* tail = ws + szslots;
    wk.err = reinterpret_cast<int*>(tail);
    wk.first_zero = reinterpret_cast<i64*>(tail + 64);
    int* ts_ws = tslu ? reinterpret_cast<int*>(tail + 256) : nullptr;
    unsigned* bmax_main = reinterpret_cast<unsigned*>(tail + 128);   // low-precision scaling
    unsigned* bmax_side = reinterpret_cast<unsigned*>(tail + 192);
    const bool lowp = sizeof(T) == 4 && lowprec != TMVX_LOWPREC_F32;
    // Trailing update A22 -= A21 U12 (K = JE - J0): FP32/FP64 MFMA, or
    // FP16/BF16 MFMA with the U12 block scaled per block row.
    auto trailing = [&](i64 rows, i64 cols, i64 K, const T* A21, const T* U12, T* A22,
                        unsigned* word, hipStream_t st) -> int {
        if constexpr (sizeof(T) == 4) {
            if (lowp)
                return gemm_lowp_launch(lowprec, rows, cols, K, -1.0f, A21, lda, U12, lda, 1.0f,
                                        A22, lda, word, st);
        }
        return gemm_launch<T>(0, 0, rows, cols, K, T(-1), A21, lda, U12, lda, T(1), A22, lda, st);
    };

    int rc = TMVX_OK;
    {
        const i64 cnt = n > nslot ? n : nslot;
        hipLaunchKernelGGL(getrf_init_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256),
                           0, s, wk.slots, nslot, wk.err, wk.first_zero, n, perm);
        rc = hip_status(hipGetLastError());
    }
    // Interchanges of panel [j0, j0+jb) on columns [a0,a1) u [b0,b1).
    auto swap_cols = [&](i64 j0, int jb, i64 a0, i64 a1, i64 b0, i64 b1, i64* pm,
                         hipStream_t st) -> int {
        const i64 cols = (a1 - a0) + (b1 - b0);
        if (cols <= 0 && !pm) return TMVX_OK;
        const int cpw = 4;
        const i64 waves = (cols + cpw - 1) / cpw;
        const unsigned grid = (unsigned)(waves > 0 ? (waves + 3) / 4 : 1);
        hipLaunchKernelGGL((laswp_panel_kernel<T, NB>), dim3(grid), dim3(256), 0, st, A, lda, n,
                           a0, a1, b0, b1, j0, jb, piv, pm, cpw);
        return hip_status(hipGetLastError());
    };
    // Rows [j0, j0+jb) of columns [c0, c0+nc) := L(j0..)^-1 times themselves.
    auto solve_rows = [&](i64 j0, int jb, i64 c0, i64 nc, hipStream_t st) -> int {
        if (nc <= 0) return TMVX_OK;
        const int cpw = 16;
        const i64 waves = (nc + cpw - 1) / cpw;
        hipLaunchKernelGGL((trsm_u12_kernel<T, NB>), dim3((unsigned)((waves + 3) / 4)),
                           dim3(256), 0, st, A, lda, j0, jb, c0, nc, cpw);
        return hip_status(hipGetLastError());
    };
    // U12 of the outer block row [J0, JE) for columns [c0, c1): NB rows at
    // a time, each followed by the update of the block row's rows below it.
    auto block_row_u12 = [&](i64 J0, i64 JE, i64 c0, i64 c1, hipStream_t st) -> int {
        for (i64 j0 = J0; j0 < JE; j0 += NB) {
            const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
            TMVX_TRY(solve_rows(j0, jb, c0, c1 - c0, st));
            const i64 below = JE - j0 - jb;
            if (below > 0)
                TMVX_TRY(gemm_launch<T>(0, 0, below, c1 - c0, jb, T(-1), A + (j0 + jb) + j0 * lda,
                                        lda, A + j0 + c0 * lda, lda, T(1),
                                        A + (j0 + jb) + c0 * lda, lda, st));
        }
        return TMVX_OK;
    };
    // Two-level blocking: outer panels of NBO columns are factorised as
    // NB-column panels updated only inside the outer panel; the trailing
    // matrix then takes one update with K = NBO (4x the arithmetic per
    // byte of C moved compared with K = NB).
    //
    // Look-ahead: after outer panel k, the caller's stream updates only the
    // next outer panel's columns and goes on to factor it, while a second
    // stream updates the rest of the trailing matrix.  The panel chain (a
    // few workgroups, latency-bound) then overlaps the big GEMM.  The next
    // panel's interchanges outside it wait for that GEMM (it reads the L
    // and U blocks the interchanges would move).
    constexpr i64 NBO = 4 * NB;
    hipStream_t side = nullptr;
    hipEvent_t ev_main = nullptr, ev_side = nullptr;
    bool side_pending = false;
    for (i64 J0 = 0; J0 < n && rc == TMVX_OK; J0 += NBO) {
        const i64 JE = (n - J0) < NBO ? n : J0 + NBO;      // outer panel [J0, JE)
        for (i64 j0 = J0; j0 < JE && rc == TMVX_OK; j0 += NB) {
            const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
            if constexpr (sizeof(T) == 4) {
                if (tslu) {
                    // tournament pivots, their interchanges on the panel's
                    // own columns and the panel's L and U
                    if ((rc = tslu_panel<T>(A, lda, n, j0, jb, piv, wk.first_zero, ts_ws, s)) !=
                        TMVX_OK)
                        break;
                }
            }
            if (!tslu) {
                const int G = (int)((n - j0 + NTH - 1) / NTH);
                hipLaunchKernelGGL((panel_kernel<T, NB, NTH>), dim3(G), dim3(NTH), 0, s,
                                   A, lda, n, j0, jb, piv, wk);
                if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
            }
            // inside the outer panel only; the rest is deferred
            if ((rc = swap_cols(j0, jb, J0, j0, j0 + jb, JE, perm, s)) != TMVX_OK) break;
            const i64 nr = JE - j0 - jb;
            if (nr > 0) {
                if ((rc = solve_rows(j0, jb, j0 + jb, nr, s)) != TMVX_OK) break;
                rc = gemm_launch<T>(0, 0, n - j0 - jb, nr, jb, T(-1), A + (j0 + jb) + j0 * lda,
                                    lda, A + j0 + (j0 + jb) * lda, lda, T(1),
                                    A + (j0 + jb) + (j0 + jb) * lda, lda, s);
            }
        }
        if (rc != TMVX_OK) break;
        if (side_pending) {
            if (hipStreamWaitEvent(s, ev_side, 0) != hipSuccess) { rc = TMVX_ERR_HIP; break; }
            side_pending = false;
        }
        // Deferred interchanges on the columns outside the outer panel.  With
        // a side stream to come, the caller's stream swaps only the next
        // outer panel's columns; the side stream swaps the rest (the left
        // L blocks and the far trailing columns) before its own work.
        const i64 nr = n - JE;
        const i64 JE2 = (n - JE) < NBO ? n : JE + NBO;      // next outer panel [JE, JE2)
        const bool has_side = nr > 0 && JE2 < n;
        for (i64 j0 = J0; j0 < JE && rc == TMVX_OK; j0 += NB) {
            const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
            rc = has_side ? swap_cols(j0, jb, JE, JE2, JE2, JE2, nullptr, s)
                          : swap_cols(j0, jb, 0, J0, JE, n, nullptr, s);
        }
        if (nr <= 0 || rc != TMVX_OK) continue;
        // next panel's columns on the caller's stream
        if ((rc = block_row_u12(J0, JE, JE, JE2, s)) != TMVX_OK) break;
        rc = trailing(nr, JE2 - JE, JE - J0, A + JE + J0 * lda, A + J0 + JE * lda,
                      A + JE + JE * lda, bmax_main, s);
        if (rc != TMVX_OK || JE2 >= n) continue;
        // the rest of the trailing matrix on the side stream
        if (!side) {
            if (hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ev_main, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&ev_side, hipEventDisableTiming) != hipSuccess) {
                rc = TMVX_ERR_HIP;
                break;
            }
        }
        if (hipEventRecord(ev_main, s) != hipSuccess ||
            hipStreamWaitEvent(side, ev_main, 0) != hipSuccess) {
            rc = TMVX_ERR_HIP;
            break;
        }
        for (i64 j0 = J0; j0 < JE && rc == TMVX_OK; j0 += NB) {
            const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
            rc = swap_cols(j0, jb, 0, J0, JE2, n, nullptr, side);
        }
        if (rc != TMVX_OK) break;
        if ((rc = block_row_u12(J0, JE, JE2, n, side)) != TMVX_OK) break;
        rc = trailing(nr, n - JE2, JE - J0, A + JE + J0 * lda, A + J0 + JE2 * lda,
                      A + JE + JE2 * lda, bmax_side, side);
        if (rc != TMVX_OK) break;
        if (hipEventRecord(ev_side, side) != hipSuccess) { rc = TMVX_ERR_HIP; break; }
        side_pending = true;
    }
    // whatever happened above, the caller's stream waits for the side stream
    if (side) {
        if (hipEventRecord(ev_side, side) != hipSuccess ||
            hipStreamWaitEvent(s, ev_side, 0) != hipSuccess)
            rc = rc == TMVX_OK ? TMVX_ERR_HIP : rc;
    }
    int herr = 0;
    i64 fz = n + 1;
    if (hipMemcpyAsync(&herr, wk.err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&fz, wk.first_zero, sizeof(i64), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    hipFreeAsync(ws, s);
    const bool sync_ok = hipStreamSynchronize(s) == hipSuccess;
    if (side) {
        hipEventDestroy(ev_main);
        hipEventDestroy(ev_side);
        hipStreamDestroy(side);
    }
    if (!sync_ok) return TMVX_ERR_HIP;
    if (rc != TMVX_OK) return rc;
    if (herr) return TMVX_ERR_TIMEOUT;
    *info = fz <= n ? fz : 0;
    return TMVX_OK;
}

// Interchange list -> composed permutation.  Step k swaps positions k and
// piv[k] >= k, so position k is final after step k.  With the positions'
// contents in LDS (n <= PERM_LDS_MAX), one wave takes the steps 64 at a
// time: lane i holds the content of position k0 + i and that of its target
// piv[k0 + i] when the target lies past the group (lanes sharing a target
// share the lowest such lane's copy); the 64 swaps run on registers
// (readlane and selects), the group's positions are then final and written
// out, and the targets past the group go back to LDS.  Larger n, or a list
// with some piv[k] < k (not an LU interchange list): one thread, in order.
constexpr int PERM_LDS_MAX = 32768;

__global__ void __launch_bounds__(1024)
piv_to_perm_kernel(i64 n, const i64* __restrict__ piv, i64* __restrict__ perm)
{
    __shared__ int sp[PERM_LDS_MAX];
    __shared__ int s_back;               // some piv[k] < k: steps in order, one thread
    const bool lds = n <= PERM_LDS_MAX;
    if (threadIdx.x == 0) s_back = 0;
    __syncthreads();
    for (i64 i = threadIdx.x; i < n; i += blockDim.x) {
        if (lds) sp[i] = (int)i; else perm[i] = i;
        const i64 p = piv[i];
        if (p >= 0 && p < i) s_back = 1;
    }
    __syncthreads();
    if (!lds || s_back) {
        if (threadIdx.x == 0) {
            for (i64 k = 0; k < n; ++k) {
                const i64 p = piv[k];
                if (p == k || p < 0 || p >= n) continue;
                if (lds) { const int x = sp[k]; sp[k] = sp[p]; sp[p] = x; }
                else { const i64 x = perm[k]; perm[k] = perm[p]; perm[p] = x; }
            }
        }
        __syncthreads();
        if (lds)
            for (i64 i = threadIdx.x; i < n; i += blockDim.x) perm[i] = sp[i];
        return;
    }
    if (threadIdx.x >= 64) return;
    const int lane = threadIdx.x;
    for (i64 k0 = 0; k0 < n; k0 += 64) {
        const int cnt = (int)((n - k0) < 64 ? (n - k0) : 64);
        const i64 k = k0 + lane;
        const bool act = lane < cnt;
        i64 pk = act ? piv[k] : k;
        if (pk < k || pk >= n) pk = k;                    // out of range: no swap
        const int p = (int)pk;
        const bool inside = pk < k0 + 64;
        int cv = act ? sp[k] : 0;                          // content of position k
        int ov = (act && !inside) ? sp[p] : 0;             // content of target p
        // owner of an outside target: the lowest lane with the same target
        int owner = lane;
        for (int j = 63; j >= 0; --j) {
            const int pj = __builtin_amdgcn_readlane(p, j);
            const int oj = __builtin_amdgcn_readlane((int)(act && !inside), j);
            if (oj && act && !inside && pj == p) owner = j;
        }
        for (int i = 0; i < cnt; ++i) {
            const int pi = __builtin_amdgcn_readlane(p, i);
            const int a = __builtin_amdgcn_readlane(cv, i);
            if (pi < k0 + 64) {                            // both positions in the group
                const int bl = pi - (int)k0;
                const int b = __builtin_amdgcn_readlane(cv, bl);
                cv = lane == i ? b : cv;
                cv = lane == bl ? a : cv;
            } else {
                const int o = __builtin_amdgcn_readlane(owner, i);
                const int b = __builtin_amdgcn_readlane(ov, o);
                cv = lane == i ? b : cv;
                ov = lane == o ? a : ov;
            }
        }
        if (act) perm[k] = cv;
        if (act && !inside && owner == lane) sp[p] = ov;
    }
}

int piv_to_perm(i64 n, const i64* piv, i64* perm, hipStream_t s)
{
    if (n <= 0) return TMVX_OK;
    hipLaunchKernelGGL(piv_to_perm_kernel, dim3(1), dim3(1024), 0, s, n, piv, perm);
    return hip_status(hipGetLastError());
}

template int getrf_impl<double>(i64, double*, i64, i64*, i64*, i64*, hipStream_t, int, int);
template int getrf_impl<float>(i64, float*, i64, i64*, i64*, i64*, hipStream_t, int, int);

} // namespace tmvx

#ifdef TMVX_PANEL_TRACE
extern "C" int tmvx_panel_trace_set(unsigned long long* buf)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(tmvx::g_panel_trace), &buf, sizeof(buf)) == hipSuccess
               ? 0 : -3;
}
#endif

extern "C" {

int tmvx_getrf_f64(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return tmvx::getrf_impl<double>(n, A, lda, piv, nullptr, info, (hipStream_t)stream); }
int tmvx_getrf_f32(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return tmvx::getrf_impl<float>(n, A, lda, piv, nullptr, info, (hipStream_t)stream); }

int tmvx_getrf_f32_tournament(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info,
                              void* stream)
{
    return tmvx::getrf_impl<float>(n, A, lda, piv, nullptr, info, (hipStream_t)stream,
                                   TMVX_LOWPREC_F32, 1);
}

} // extern "C"
