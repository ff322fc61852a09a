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
#include <mutex>
#include <vector>

namespace tmvx {

constexpr int PANEL_NB = 64;        // panel width (columns)
// FP32 panels may be 128 columns wide (-DTMVX_F32_NB=128: 212 VGPRs, two
// waves per SIMD, no spills), which halves the launches on the critical
// path; measured on MI355X (FP32 LU, n = 16384): 99.9 ms against 87.8 ms
// with 64 columns -- each column step of the wider panel costs more than
// the launches save -- so 64 is the default.
#ifndef TMVX_F32_NB
#define TMVX_F32_NB 64
#endif
// The next outer panel's U12 on the caller's stream: in the fused launch
// (1) or on the trsm / GEMM kernels after a fused interchange launch (0).
#ifndef TMVX_K2_FUSED
#define TMVX_K2_FUSED 1
#endif
// Outer panels are NBO_MUL panels wide (trailing updates with K = NBO_MUL NB).
#ifndef TMVX_NBO_MUL_F64
#define TMVX_NBO_MUL_F64 4
#endif
#ifndef TMVX_NBO_MUL
#define TMVX_NBO_MUL 4
#endif
constexpr int PANEL_BB = 8;         // columns per delayed-update block
// Rows per workgroup (one per thread); a build option for timing the
// choice: -DTMVX_PANEL_THREADS=256.
#ifndef TMVX_PANEL_THREADS
#define TMVX_PANEL_THREADS 512
#endif
constexpr int PANEL_THREADS = TMVX_PANEL_THREADS;
constexpr int PANEL_GMAX = 512;     // workgroups per panel (n <= 262144)
// Bounded waits: a wait that sees no progress for this long (wall clock,
// hipDeviceAttributeWallClockRate = 100 MHz on MI355X) raises the error flag.
constexpr long long WAIT_TICKS = 200000000ll;   // 2 s

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
    unsigned long long* cslots;  // [2][ColSlot::SIZE][G] per-column candidates
    unsigned long long* bslots;  // [2][BB][NB * WPV] per-block pivot rows
    int* err;                    // timeout flag
    i64* first_zero;             // 1-based first zero pivot (min), n+1 if none
};

// Per-column exchange slot of one workgroup, in tagged 64-bit words:
// (column tag << 32) | 32 payload bits.  A double takes two words (low,
// high half), a float one.  Each word is written and read as one 64-bit
// atomic, so a reader that sees the current tag in a word also sees that
// word's payload: no separate flag and no ordering fence is needed.
template <typename T, int BB> struct ColSlot {
    static constexpr int WPV = sizeof(T) / 4;   // words per value
    static constexpr int VAL = 0;               // candidate |value|
    static constexpr int POS = WPV;             // candidate row position
    static constexpr int ROW = WPV + 1;         // candidate's BB block values
    static constexpr int SIZE = ROW + BB * WPV;
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

// Timestamps of the panel loop, only in a build with -DTMVX_PANEL_TRACE,
// for the panel starting at column g_panel_trace_j0: per column, thread 0
// of every workgroup records the wall clock (100 MHz, common to all CUs)
// after the barrier, at publication, when every slot has arrived and after
// the update: g_panel_trace[(k * G + w) * 4 + phase].
#ifdef TMVX_PANEL_TRACE
__device__ unsigned long long* g_panel_trace;
__device__ long long g_panel_trace_j0;
#define PANEL_TRACE_INIT()                                                     \
    unsigned long long* const trace_buf =                                      \
        (t == 0 && j0 == g_panel_trace_j0) ? g_panel_trace : nullptr
#define PANEL_TRACE(k, ph)                                                     \
    do {                                                                       \
        if (trace_buf) trace_buf[((k) * G + w) * 4 + (ph)] = wall_clock64();   \
    } while (0)
#else
#define PANEL_TRACE_INIT() do {} while (0)
#define PANEL_TRACE(k, ph) do {} while (0)
#endif

// One panel [j0, j0+jb) x rows [j0, n): GEPP with TMV's pivot rules.
//
// Rows never move while the panel is factored: each thread keeps its row
// in registers together with the row's current position in the swapped
// order (pos) and whether it has become a pivot row (used).  At step k the
// candidates are the unused rows (exactly those at positions >= k); the
// winner p (largest |a|, then lowest position: TMV's strict '>') takes
// position k and the row that was at k takes p's old position, which is
// all a row interchange changes.  At the end every row is written to its
// final position, so the interchanges on the panel's own columns cost
// nothing extra.
//
// Updates are delayed by blocks of BB columns: a step updates only the
// current block's columns of the unused rows; at the end of a block every
// workgroup receives the block's BB pivot rows, forms U12 = L11^-1 A12 for
// the columns to the right, and updates its rows with one rank-BB product.
// The per-column exchange therefore carries only (|value|, position, BB
// block values) per workgroup.  Operation order per element is that of the
// unblocked right-looking elimination, so the factors do not depend on BB.
//
// Registers hold the row rotated: by whole blocks at block ends (before
// block kb, row[c] is panel column (kb*BB + c) mod NB for c >= BB), and
// within the block by one column per step (the current column is always
// row[0]), so the column loop is one copy of code with static register
// indices.
//
// Per column: each wave's argmax goes to LDS (double-buffered by column
// parity), one barrier, wave 0 reduces them, publishes the workgroup's
// candidate, receives every workgroup's slot (lane q holds slot q), picks
// the winner and puts it in LDS; a second barrier hands it to every wave.
template <typename T, int NB, int BB, int NTH, bool SOLO>
__global__ void __launch_bounds__(NTH)
panel_kernel(T* __restrict__ A, i64 lda, i64 n, i64 j0, int jb, i64* __restrict__ piv,
             PanelWork wk)
{
    typedef ColSlot<T, BB> CS;
    constexpr int WPV = CS::WPV;
    constexpr int NW = NTH / 64;
    constexpr int NBLK = NB / BB;
    constexpr int RW = NB * WPV;                 // words of one staged row
    constexpr int BW = BB * WPV;                 // words of a candidate's block values
    static_assert(NB % BB == 0 && BB % 4 == 0 && NW <= 8 && CS::SIZE <= 64, "panel shape");
    const int G = gridDim.x, w = blockIdx.x, t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const i64 m = n - j0;
    const int r = w * NTH + t;                   // panel-relative row
    const bool own = r < m;

    __shared__ T s_wv[2][NW];                    // per-wave candidates (column parity)
    __shared__ int s_wi[2][NW];
    __shared__ __attribute__((aligned(16))) T s_wrow[2][NW][BB];   // their block values
    __shared__ __attribute__((aligned(16))) T s_pv[BB];            // the pivot row's block values
    __shared__ int s_p, s_zero, s_fail;
    // the panel's interchange list, written out at the end: a global store
    // in the column loop would make the next barrier wait for it
    __shared__ int s_piv[NB];
    __shared__ int s_own[BB];                    // block pivot i is held here
    // block pivot rows (block order), as 32-bit words: staged by their
    // owners or received; U12 is solved in place in columns >= BB
    __shared__ __attribute__((aligned(16))) unsigned s_rw[BB * RW];
    T (*s_rows)[NB] = reinterpret_cast<T (*)[NB]>(s_rw);

    // Panel waves go first on SIMDs they share with other kernels' waves.
    __builtin_amdgcn_s_setprio(3);
    PANEL_TRACE_INIT();

    T row[NB];
    {
        const T* gp = A + (j0 + r) + j0 * lda;
#pragma unroll
        for (int c = 0; c < NB; ++c) row[c] = (own && c < jb) ? gp[(i64)c * lda] : T(0);
    }
    int pos = r;                  // current position (panel-relative)
    bool used = !own;             // a pivot row (or no row at all)
    int mypiv = -1;               // index of this row among the current block's pivots
    bool dfr = false;             // takes the previous block's deferred update chunks
    int nchunk = 0;               // deferred chunks pending (uniform)
    if (t == 0) s_fail = 0;
    if (t < BB) s_own[t] = 0;
    bool alive = __hip_atomic_load(wk.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    int nrot = 0;                 // whole-block rotations so far

    auto rotate_blocks = [&]() __attribute__((always_inline)) {
        T first[BB];
#pragma unroll
        for (int q = 0; q < BB; ++q) first[q] = row[q];
#pragma unroll
        for (int c = 0; c < NB - BB; ++c) row[c] = row[c + BB];
#pragma unroll
        for (int q = 0; q < BB; ++q) row[NB - BB + q] = first[q];
        ++nrot;
    };
    auto rotate_in_block = [&]() __attribute__((always_inline)) {
        const T first = row[0];
#pragma unroll
        for (int q = 0; q < BB - 1; ++q) row[q] = row[q + 1];
        row[BB - 1] = first;
    };
    // Deferred share of the previous block's rank-BB update: chunk j (the
    // BB columns now at row[j*BB ..]) with the block's multipliers (now at
    // row[NB-BB ..]) and its U12 rows (still in s_rows, columns (j+1)*BB ..).
    // Applied while the workgroup waits for the exchange, chunk j at column
    // step (j - 1) mod BB (one chunk per step for NB = 64, up to two for
    // NB = 128); all are done before the block's own end.
    auto deferred_chunk = [&](int kk) __attribute__((always_inline)) {
        if (!dfr) return;
        typedef typename Vec16<T>::type VT;
        constexpr int VN = Vec16<T>::n;
        static_assert(NBLK - 2 <= 2 * BB, "two chunks per step at most");
        static_for<1, NBLK - 1>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if ((j - 1) % BB != kk || j > nchunk) return;             // uniform
            static_for<0, BB>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const T li = row[NB - BB + i];
                static_for<0, BB / VN>([&](auto qc) {
                    constexpr int q = decltype(qc)::value * VN;
                    const VT u = *reinterpret_cast<const VT*>(&s_rows[i][(j + 1) * BB + q]);
                    static_for<0, VN>([&](auto uc) {
                        constexpr int x = decltype(uc)::value;
                        row[j * BB + q + x] -= li * u[x];
                    });
                });
            });
            // keeps the per-chunk branches apart: merged, they would index
            // row[] with the runtime chunk number and move it to memory
            __asm__ volatile("" ::: "memory");
        });
    };
    // Bounded wait bookkeeping: false once the wait has to give up --
    // another workgroup raised the error flag, or no progress for WAIT_TICKS.
    auto still_waiting = [&](unsigned& spins, long long& t0) __attribute__((always_inline)) -> bool {
        if ((++spins & 63) != 0) return true;
        if (__hip_atomic_load(wk.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        const long long now = wall_clock64();
        if (t0 == 0) t0 = now;
        if (now - t0 > WAIT_TICKS) {
            __hip_atomic_store(wk.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        return true;
    };

    const int nblk = (jb + BB - 1) / BB;
    for (int kb = 0; kb < nblk && alive; ++kb) {
        const int c0 = kb * BB;
        const int bcnt = (jb - c0) < BB ? (jb - c0) : BB;
        int kk = 0;
        for (; kk < bcnt; ++kk) {
            const int k = c0 + kk;
            const int par = k & 1;
            const unsigned tag32 = (unsigned)(j0 + k + 1);
            const unsigned long long tag = (unsigned long long)tag32 << 32;
            // -- wave argmax over the unused rows ---------------------------
            T v;
            int vi;
            lanes_argmax<64>(used ? T(-1) : fabs(row[0]), pos, v, vi);
            if (lane == 0) { s_wv[par][wave] = v; s_wi[par][wave] = vi; }
            if (!used && pos == vi) {
                typedef typename Vec16<T>::type VT;
                constexpr int VN = Vec16<T>::n;
                static_for<0, BB / VN>([&](auto qc) {
                    constexpr int q = decltype(qc)::value * VN;
                    VT x;
                    static_for<0, VN>([&](auto uc) { x[decltype(uc)::value] = row[q + decltype(uc)::value]; });
                    *reinterpret_cast<VT*>(&s_wrow[par][wave][q]) = x;
                });
            }
            __syncthreads();
            PANEL_TRACE(k, 0);
            if (wave != 0) deferred_chunk(kk);     // inside the exchange wait
            if (wave == 0) {
                // -- this workgroup's candidate, published as tagged words;
                //    word f of workgroup q lives at cs[f * G + q] ----------
                unsigned long long* cs = wk.cslots + (i64)par * CS::SIZE * G;
                const int myi = lane < NW ? s_wi[par][lane] : INT_MAX;
                T bv;
                int bi;
                lanes_argmax<8>(lane < NW ? s_wv[par][lane] : T(-1), myi, bv, bi);
                const unsigned long long hw = __ballot(lane < NW && myi == bi);
                const int ww = hw ? (int)__builtin_ctzll(hw) : 0;
                if (!SOLO && lane < CS::SIZE) {
                    const int e = lane >= CS::ROW ? lane - CS::ROW : 0;
                    const unsigned rowpay = half_bits(s_wrow[par][ww][e / WPV], e % WPV);
                    const unsigned pay = lane < CS::POS ? half_bits(bv, lane)
                                       : lane == CS::POS ? (unsigned)bi : rowpay;
                    cstore(cs + (i64)lane * G + w, tag | pay);
                }
                PANEL_TRACE(k, 1);
                deferred_chunk(kk);                // before any slot can arrive
                // -- receive every slot: lane q holds slot q ----------------
                T bestv = T(-1);
                int besti = INT_MAX;
                unsigned bpay[BW];
#pragma unroll
                for (int f = 0; f < BW; ++f) bpay[f] = 0u;
                bool ok = true;
                auto take = [&](const unsigned long long (&wv)[CS::SIZE], bool valid)
                    __attribute__((always_inline)) {
                    if (!valid) return;
                    unsigned pay[CS::SIZE];
#pragma unroll
                    for (int f = 0; f < CS::SIZE; ++f) pay[f] = (unsigned)wv[f];
                    const T qv = from_halves<T>(&pay[CS::VAL]);
                    const int qi = (int)pay[CS::POS];
                    if (better(qv, qi, bestv, besti)) {
                        bestv = qv;
                        besti = qi;
#pragma unroll
                        for (int f = 0; f < BW; ++f) bpay[f] = pay[CS::ROW + f];
                    }
                };
                auto complete = [&](const unsigned long long (&wv)[CS::SIZE], bool valid)
                    __attribute__((always_inline)) -> bool {
                    bool have = true;
                    static_for<0, CS::SIZE>([&](auto fc) {
                        constexpr int f = decltype(fc)::value;
                        have &= (unsigned)(wv[f] >> 32) == tag32;
                    });
                    return __ballot(valid && !have) == 0ull;
                };
                if constexpr (SOLO) {
                    // one workgroup holds every row: its candidate is the
                    // winner, no exchange
                    if (lane == 0) {
                        bestv = bv;
                        besti = bi;
#pragma unroll
                        for (int f = 0; f < BW; ++f)
                            bpay[f] = half_bits(s_wrow[par][ww][f / WPV], f % WPV);
                    }
                } else {
                    for (int qb = 0; qb < G && ok; qb += 64) {
                        const int q = qb + lane;
                        const bool valid = q < G;
                        const unsigned long long* src = cs + (valid ? q : qb);
                        unsigned long long wv[CS::SIZE];
                        unsigned spins = 0;
                        long long t0 = 0;
                        while (true) {
                            static_for<0, CS::SIZE>([&](auto fc) {
                                constexpr int f = decltype(fc)::value;
                                wv[f] = cload(src + (i64)f * G);
                            });
                            if (complete(wv, valid)) break;
                            if (!still_waiting(spins, t0)) { ok = false; break; }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        if (ok) take(wv, valid);
                    }
                }
                PANEL_TRACE(k, 2);
                if (ok) {
                    // -- the global winner (the same in every workgroup) ----
                    T gv;
                    int gi;
                    lanes_argmax<64>(bestv, besti, gv, gi);
                    const unsigned long long hg = __ballot(besti == gi);
                    const int wl = hg ? (int)__builtin_ctzll(hg) : 0;
                    const bool zero = !(gv >= tiny<T>());              // TMV_Underflow
                    const int p = zero ? k : gi;
                    // the winner's block values, written by the lane that
                    // holds them (readlanes would each wait for an SGPR)
                    if (lane == wl) {
                        static_for<0, BB>([&](auto jc) {
                            constexpr int j = decltype(jc)::value;
                            s_pv[j] = from_halves<T>(&bpay[j * WPV]);
                        });
                    }
                    if (lane == 0) {
                        s_zero = zero ? 1 : 0;
                        s_p = p;
                        if (w == 0) {
                            s_piv[k] = p;
                            if (zero)
                                atomicMin(reinterpret_cast<unsigned long long*>(wk.first_zero),
                                          (unsigned long long)(j0 + k + 1));
                        }
                    }
                } else if (lane == 0) {
                    s_fail = 1;
                }
            }
            __syncthreads();
            if (s_fail) { alive = false; break; }                    // uniform after the barrier
            // -- interchange (positions only) and the block's update ---------
            const int p = s_p;
            if (s_zero) {
                // TMV: zero the column from the diagonal down, interchange
                // with itself; the row at position k is the pivot row.
                if (!used) row[0] = T(0);
                if (!used && pos == k) { used = true; mypiv = kk; }
            } else {
                if (!used && pos == p) { used = true; mypiv = kk; pos = k; }
                else if (!used && pos == k) pos = p;
                if (!used) {
                    typedef typename Vec16<T>::type VT;
                    constexpr int VN = Vec16<T>::n;
                    T pv[BB];
                    static_for<0, BB / VN>([&](auto qc) {
                        constexpr int q = decltype(qc)::value * VN;
                        const VT x = *reinterpret_cast<const VT*>(&s_pv[q]);
                        static_for<0, VN>([&](auto uc) { pv[q + decltype(uc)::value] = x[decltype(uc)::value]; });
                    });
                    const T l = row[0] * (T(1) / pv[0]);
                    row[0] = l;
                    const int lim = BB - kk;       // block columns after this one
#pragma unroll
                    for (int c = 1; c < BB; ++c)
                        if (c < lim) row[c] -= l * pv[c];
                }
            }
            rotate_in_block();
            PANEL_TRACE(k, 3);
        }
        if (!alive) break;
        // complete the in-block rotation of a partial block
        for (; kk < BB; ++kk) rotate_in_block();
        if (c0 + BB >= jb) break;
        // -- block end: the BB pivot rows, U12, rank-BB update ---------------
        if (mypiv >= 0) {
            s_own[mypiv] = 1;
            typedef typename Vec16<T>::type VT;
            constexpr int VN = Vec16<T>::n;
            static_for<0, NB / VN>([&](auto qc) {
                constexpr int q = decltype(qc)::value * VN;
                VT x;
                static_for<0, VN>([&](auto uc) { x[decltype(uc)::value] = row[q + decltype(uc)::value]; });
                *reinterpret_cast<VT*>(&s_rows[mypiv][q]) = x;
            });
        }
        __syncthreads();
        {
            const int bpar = kb & 1;
            const unsigned long long btag = (unsigned long long)(j0 + c0 + 1) << 32;
            unsigned long long* bs = wk.bslots + (i64)bpar * BB * RW;
            if (!SOLO)                       // a single workgroup owns every pivot row
                for (int e = t; e < BB * RW; e += NTH)
                    if (s_own[e / RW]) cstore(bs + e, btag | s_rw[e]);
            // words held elsewhere: all of a thread's loads in flight at once
            constexpr int RPT = (BB * RW + NTH - 1) / NTH;
            unsigned pend = 0;
            static_for<0, RPT>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int e = t + i * NTH;
                if (e < BB * RW && !s_own[e / RW]) pend |= 1u << i;
            });
            unsigned spins = 0;
            long long t0 = 0;
            while (pend) {
                unsigned long long x[RPT];
                static_for<0, RPT>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    const int e = t + i * NTH;
                    x[i] = cload(bs + (e < BB * RW ? e : 0));
                });
                static_for<0, RPT>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    if (((pend >> i) & 1u) && (x[i] & ~0xffffffffull) == btag) {
                        pend &= ~(1u << i);
                        s_rw[t + i * NTH] = (unsigned)x[i];
                    }
                });
                if (!pend) break;
                if (!still_waiting(spins, t0)) { s_fail = 1; break; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (s_fail) { alive = false; break; }
        if (t < BB) s_own[t] = 0;
        // U12 = L11^-1 A12 in place: thread c solves column c of the rows
        // (L11(i, j) = s_rows[i][j], j < i, unit diagonal).
        if (t >= BB && t < NB - c0) {
            T u[BB];
            static_for<0, BB>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                T x = s_rows[i][t];
                static_for<0, i>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    x -= s_rows[i][j] * u[j];
                });
                u[i] = x;
                s_rows[i][t] = x;
            });
        }
        __syncthreads();
        dfr = false;
        if (mypiv >= 0) {
            // this block's pivot rows take their U12 row
            static_for<1, NBLK>([&](auto chc) {
                constexpr int ch = decltype(chc)::value;
                if (ch * BB < NB - c0) {
#pragma unroll
                    for (int q = 0; q < BB; ++q) row[ch * BB + q] = s_rows[mypiv][ch * BB + q];
                }
            });
            mypiv = -1;
        } else if (!used) {
            // rank-BB update of the next block's columns now; the columns
            // beyond it follow chunk by chunk during the next block's steps
            typedef typename Vec16<T>::type VT;
            constexpr int VN = Vec16<T>::n;
            static_for<0, BB>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const T li = row[i];
                static_for<0, BB / VN>([&](auto qc) {
                    constexpr int q = BB + decltype(qc)::value * VN;
                    const VT u = *reinterpret_cast<const VT*>(&s_rows[i][q]);
                    static_for<0, VN>([&](auto uc) {
                        constexpr int x = decltype(uc)::value;
                        row[q + x] -= li * u[x];
                    });
                });
            });
            dfr = true;
        }
        rotate_blocks();
        nchunk = nblk - kb - 2 > 0 ? nblk - kb - 2 : 0;
        // s_rows (U12) is read by the deferred chunks of the next block and
        // rewritten only at its end, after every chunk has been applied
    }
    // Restore column order (the rotation is by whole blocks).
    while ((nrot * BB) % NB != 0) rotate_blocks();
    if (!alive) return;
    if (w == 0 && t < jb) piv[j0 + t] = j0 + s_piv[t];
    if (!own) return;
    T* dst = A + (j0 + pos) + j0 * lda;
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < jb) dst[(i64)c * lda] = row[c];
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
    static_assert(NB == 64 || NB == 128, "one or two panel columns per lane");
    constexpr int NPL = NB / 64;
    const int lane = threadIdx.x & 63;
    // The state lives in P (LDS): src by panel row, osrc by outside slot,
    // read by the whole wave as broadcasts and written by lane 0 (the wave's
    // LDS accesses are in order, so each step sees the previous one's
    // stores); odst stays in registers for the ballot search.  (A readlane
    // per entry would wait for its SGPR on every step.)
    long long odst[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
        P.src[lane + 64 * q] = j0 + lane + 64 * q;
        P.osrc[lane + 64 * q] = -1;
        odst[q] = -1;
    }
    __builtin_amdgcn_wave_barrier();
    int nout = 0;
    for (int k = 0; k < jb; ++k) {
        const i64 p = lpiv[k];
        if (p <= j0 + k || p >= nrows) continue;            // uniform
        if (p < j0 + jb) {
            const int kp = (int)(p - j0);
            const i64 a = P.src[k], b = P.src[kp];
            if (lane == 0) { P.src[k] = b; P.src[kp] = a; }
        } else {
            int q = -1;
#pragma unroll
            for (int c = NPL - 1; c >= 0; --c) {
                const unsigned long long hit =
                    __ballot(lane + 64 * c < nout && odst[c] == (long long)p);
                if (hit) q = 64 * c + __ffsll((long long)hit) - 1;
            }
            if (q < 0) {
                q = nout++;
#pragma unroll
                for (int c = 0; c < NPL; ++c)
                    if (lane + 64 * c == q) odst[c] = p;
                if (lane == 0) P.osrc[q] = p;
            }
            const i64 a = P.src[k], b = P.osrc[q];
            if (lane == 0) { P.src[k] = b; P.osrc[q] = a; }
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int q = 0; q < NPL; ++q) P.odst[lane + 64 * q] = odst[q];
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
    constexpr int NPL = NB / 64;
    __shared__ PanelPerm<NB> P;
    __shared__ i64 lpiv[NB];
    if (threadIdx.x < jb) lpiv[threadIdx.x] = piv[j0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < 64) panel_perm_wave<NB>(P, lpiv, j0, jb, nrows);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int nout = P.nout;
    if (perm != nullptr && blockIdx.x == 0 && threadIdx.x < 64) {
        i64 v1[NPL], v2[NPL];
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
            const int x = lane + 64 * q;
            v1[q] = x < jb ? perm[P.src[x]] : 0;
            v2[q] = x < nout ? perm[P.osrc[x]] : 0;
        }
        __builtin_amdgcn_s_waitcnt(0);     // every load returned before any store
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
            const int x = lane + 64 * q;
            if (x < jb) perm[j0 + x] = v1[q];
            if (x < nout) perm[P.odst[x]] = v2[q];
        }
    }
    const i64 nleft = a1 - a0, nright = b1 - b0, total = nleft + nright;
    const i64 wave = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int cq = 0; cq < cols_per_wave; ++cq) {
        const i64 idx = wave * cols_per_wave + cq;
        if (idx >= total) break;
        const i64 col = idx < nleft ? a0 + idx : b0 + (idx - nleft);
        T* a = A + col * lda;
        T v1[NPL], v2[NPL];
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
            const int x = lane + 64 * q;
            v1[q] = x < jb ? a[P.src[x]] : T(0);
            v2[q] = x < nout ? a[P.osrc[x]] : T(0);
        }
        __builtin_amdgcn_s_waitcnt(0);     // every load returned before any store
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
            const int x = lane + 64 * q;
            if (x < jb) a[j0 + x] = v1[q];
            if (x < nout) a[P.odst[x]] = v2[q];
        }
    }
}

// U12 = L11^-1 A12, L11 = unit-lower A[j0:j0+jb, j0:j0+jb],
// A12 = A[j0:j0+jb, c0 : c0+ncols].  Lane i of a wave owns row i:
// it keeps row i of L11 in registers and, column by column, forward-
// substitutes with x_k broadcast by readlane.  Each wave solves CW columns
// at once (independent chains interleave) and walks `cpw` columns.  A
// 128-row panel is two 64-row halves: solve the top, subtract its product
// with the lower-left block from the bottom, solve the bottom (lane i then
// owns rows i and 64 + i).
constexpr int U12_CW = 4;

template <typename T, int NB>
__global__ void __launch_bounds__(256)
trsm_u12_kernel(T* __restrict__ A, i64 lda, i64 j0, int jb, i64 cbeg, i64 ncols, int cpw)
{
    static_assert(NB == 64 || NB == 128, "one or two panel rows per lane");
    const int lane = threadIdx.x & 63;
    const i64 wave = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const i64 c0 = wave * cpw;
    if (c0 >= ncols) return;
    const i64 c1 = c0 + cpw < ncols ? c0 + cpw : ncols;
    const int jb1 = jb < 64 ? jb : 64;           // rows of the top half
    T Lrow[64];
    const T* arow = A + (j0 + lane) + j0 * lda;
    static_for<0, 64>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        Lrow[k] = (lane < jb1 && k < lane) ? arow[(i64)k * lda] : T(0);
    });
    T Lbl[NB == 128 ? 64 : 1], Lbr[NB == 128 ? 64 : 1];
    const bool two = NB == 128 && jb > 64;
    const int r2 = 64 + lane;                    // bottom-half row of this lane
    if constexpr (NB == 128) {
        const T* brow = A + (j0 + r2) + j0 * lda;
        static_for<0, 64>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            Lbl[k] = (two && r2 < jb) ? brow[(i64)k * lda] : T(0);
            Lbr[k] = (two && r2 < jb && k < lane) ? brow[(i64)(64 + k) * lda] : T(0);
        });
    }
    T* bcol = A + (j0 + lane) + cbeg * lda;
    for (i64 c = c0; c < c1; c += U12_CW) {
        T x[U12_CW], y[U12_CW];
#pragma unroll
        for (int q = 0; q < U12_CW; ++q) {
            x[q] = (lane < jb1 && c + q < c1) ? bcol[(c + q) * lda] : T(0);
            y[q] = (two && r2 < jb && c + q < c1) ? bcol[64 + (c + q) * lda] : T(0);
        }
        static_for<0, 63>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
#pragma unroll
            for (int q = 0; q < U12_CW; ++q) {
                const T xk = readlane(x[q], k);
                if (lane > k) x[q] -= Lrow[k] * xk;
            }
        });
        if constexpr (NB == 128) {
            if (two) {
                static_for<0, 64>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
#pragma unroll
                    for (int q = 0; q < U12_CW; ++q) y[q] -= Lbl[k] * readlane(x[q], k);
                });
                static_for<0, 63>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
#pragma unroll
                    for (int q = 0; q < U12_CW; ++q) {
                        const T yk = readlane(y[q], k);
                        if (lane > k) y[q] -= Lbr[k] * yk;
                    }
                });
            }
        }
#pragma unroll
        for (int q = 0; q < U12_CW; ++q) {
            if (lane < jb1 && c + q < c1) bcol[(c + q) * lda] = x[q];
            if (two && r2 < jb && c + q < c1) bcol[64 + (c + q) * lda] = y[q];
        }
    }
}

// The interchanges piv[Jp .. Jp + rows) of up to SS_MAXP 64-column panels
// on columns [a0, a1) u [b0, b1), panel by panel; the columns in [b0, b1)
// then take X := L^-1 X on rows [Jp, Jp + rows), L the unit-lower diagonal
// block of those rows (the U12 block row, TMV's A01 /= A00.lowerTri(
// UnitDiag), src/TMV_LUDecompose.cpp:262-264).  One launch replaces the
// per-panel interchange and U12 launches.  Wave q of each workgroup turns
// panel q's interchange list into its net row movement (panel_perm_wave);
// then a wave takes its columns one at a time: all loads of a panel
// before its stores, lane i keeps row 64 q + i of panel q in registers,
// and the solve runs block by block, the L rows read from memory (L2) in
// slices, the solved blocks broadcast to the blocks below through LDS.
// `perm` (the composed permutation of the mixed-precision solver) moves
// with the rows.
constexpr int SS_CW = 1;
template <typename T, int SS_MAXP>                // SS_MAXP: panels per launch (NBO / NB)
__global__ void __launch_bounds__(256)
swap_solve_kernel(T* __restrict__ A, i64 lda, i64 nrows, i64 Jp, int rows,
                  const i64* __restrict__ piv, i64 a0, i64 a1, i64 b0, i64 b1, int cpw,
                  i64* __restrict__ perm, int solve_b)
{
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int np = (rows + 63) >> 6;
    // solved blocks of this wave's columns, read back as broadcasts by the
    // blocks below them (a readlane per value would be hoisted en masse)
    __shared__ T s_x[4][SS_MAXP - 1][64][SS_CW];
    __shared__ PanelPerm<64> P[SS_MAXP];
    __shared__ i64 lpiv[SS_MAXP * 64];
    for (int i = threadIdx.x; i < rows; i += blockDim.x) lpiv[i] = piv[Jp + i];
    __syncthreads();
    for (int q = wv; q < np; q += 4) {
        const int jbq = rows - 64 * q < 64 ? rows - 64 * q : 64;
        panel_perm_wave<64>(P[q], lpiv + 64 * q, Jp + 64 * q, jbq, nrows);
    }
    __syncthreads();
    // lane i: row 64 q + i of panel q comes from src[q]; the q-th panel's
    // i-th outside row moves from osrc[q] to odst[q] (i < nout)
    i64 src[SS_MAXP], osrc[SS_MAXP], odst[SS_MAXP];
#pragma unroll
    for (int q = 0; q < SS_MAXP; ++q) {
        const bool in = q < np && 64 * q + lane < rows;
        const bool out = q < np && lane < P[q < np ? q : 0].nout;
        src[q] = in ? P[q].src[lane] : 0;
        osrc[q] = out ? P[q].osrc[lane] : -1;
        odst[q] = out ? P[q].odst[lane] : -1;
    }
    if (perm != nullptr && blockIdx.x == 0 && threadIdx.x < 64) {
        static_for<0, SS_MAXP>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (q >= np) return;                                      // uniform
            const int i = 64 * q + lane;
            const i64 v1 = i < rows ? perm[src[q]] : 0;
            const i64 v2 = osrc[q] >= 0 ? perm[osrc[q]] : 0;
            __builtin_amdgcn_s_waitcnt(0);           // every load before any store
            if (i < rows) perm[Jp + i] = v1;
            if (osrc[q] >= 0) perm[odst[q]] = v2;
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        });
    }
    const i64 nleft = a1 - a0, nright = b1 - b0, total = nleft + nright;
    const i64 wave = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const i64 cbeg = wave * cpw, cend = cbeg + cpw < total ? cbeg + cpw : total;
    for (i64 c = cbeg, cl; c < cend; c = cl) {
        // a group of up to SS_CW columns of one range (a group stops at the
        // end of the left range)
        const bool solve = c >= nleft && solve_b;
        cl = c + SS_CW < cend ? c + SS_CW : cend;
        if (c < nleft && cl > nleft) cl = nleft;
        T* col[SS_CW];
#pragma unroll
        for (int w = 0; w < SS_CW; ++w) {
            const i64 idx = c + w < cl ? c + w : c;
            col[w] = A + (idx < nleft ? a0 + idx : b0 + (idx - nleft)) * lda;
        }
        const int ncw = (int)(cl - c);
        T x[SS_MAXP][SS_CW];
        static_for<0, SS_MAXP>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
#pragma unroll
            for (int w = 0; w < SS_CW; ++w) x[q][w] = T(0);
            if (q >= np) return;                                      // uniform
            const int i = 64 * q + lane;
            T v1[SS_CW], v2[SS_CW];
#pragma unroll
            for (int w = 0; w < SS_CW; ++w) {
                v1[w] = (w < ncw && i < rows) ? col[w][src[q]] : T(0);
                v2[w] = (w < ncw && osrc[q] >= 0) ? col[w][osrc[q]] : T(0);
            }
            __builtin_amdgcn_s_waitcnt(0);           // every load before any store
#pragma unroll
            for (int w = 0; w < SS_CW; ++w) {
                if (w < ncw && i < rows && !solve) col[w][Jp + i] = v1[w];
                if (w < ncw && osrc[q] >= 0) col[w][odst[q]] = v2[w];
                x[q][w] = v1[w];
            }
            __builtin_amdgcn_s_waitcnt(0);           // stores done before the next panel's loads
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        });
        if (!solve) continue;                        // (the loop moves c to cl)
        // X := L^-1 X, block by block (rows 64 q + lane); the empty asm
        // statements keep each block's L loads inside its own step
        static_for<0, SS_MAXP>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (q >= np) return;                                      // uniform
            const int i = 64 * q + lane;
            const T* lrow = A + (Jp + (i < rows ? i : 0)) + Jp * lda;
            // L is read in slices of 16 columns (16 loads in flight); the row
            // pointer passes through an empty asm statement that takes the
            // running x as input, so a slice's loads wait for the previous
            // slice's work instead of all being hoisted to the top
            constexpr int KS = sizeof(T) == 4 ? 64 : 32;
            static_for<0, q>([&](auto qqc) {
                constexpr int qq = decltype(qqc)::value;
                static_for<0, 64 / KS>([&](auto sc) {
                    constexpr int k0 = decltype(sc)::value * KS;
                    const T* lp = lrow;
                    __asm__ volatile("" : "+v"(lp) : "v"(x[q][0]));
                    T L[KS];
                    static_for<0, KS>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        L[k] = i < rows ? lp[(i64)(64 * qq + k0 + k) * lda] : T(0);
                    });
                    static_for<0, KS>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
#pragma unroll
                        for (int w = 0; w < SS_CW; ++w)
                            x[q][w] -= L[k] * s_x[wv][qq][k0 + k][w];   // broadcast read
                    });
                });
            });
            static_for<0, 64 / KS>([&](auto sc) {
                constexpr int k0 = decltype(sc)::value * KS;
                const T* lp = lrow;
                __asm__ volatile("" : "+v"(lp) : "v"(x[q][0]));
                T L[KS];
                static_for<0, KS>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    L[k] = (i < rows && k0 + k < lane) ? lp[(i64)(64 * q + k0 + k) * lda] : T(0);
                });
                static_for<0, KS>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    if constexpr (k0 + k < 63) {
#pragma unroll
                        for (int w = 0; w < SS_CW; ++w) {
                            const T xk = readlane(x[q][w], k0 + k);
                            if (lane > k0 + k) x[q][w] -= L[k] * xk;
                        }
                    }
                });
            });
#pragma unroll
            for (int w = 0; w < SS_CW; ++w) {
                if (w < ncw && i < rows) col[w][Jp + i] = x[q][w];
                if (q + 1 < SS_MAXP) s_x[wv][q][lane][w] = x[q][w];      // for the blocks below
            }
            __asm__ volatile("" ::: "memory");
        });
    }
}

// LU without interchanges (the mixed-precision solver after a random
// butterfly transformation; never the TMV-parity path).  The panel
// [j0, j0+jb) is two launches with no grid-wide exchange:
//
// genp_diag_kernel: the jb x jb diagonal block, one wave, lane i holding
// row i in registers.  Step k broadcasts the pivot row by readlane, scales
// column k below the diagonal by the reciprocal of the pivot and updates the
// rows below.  A zero or non-finite pivot is recorded in *first_zero
// (1-based, minimum) and its column is left unscaled and unapplied (the
// caller treats the factorisation as failed).  piv receives the identity
// interchange list.
template <typename T>
__global__ void __launch_bounds__(64)
genp_diag_kernel(T* __restrict__ A, i64 lda, i64 j0, int jb, i64* __restrict__ piv,
                 i64* __restrict__ first_zero)
{
    const int lane = threadIdx.x;
    T row[64];
    T* ap = A + (j0 + lane) + j0 * lda;
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        row[c] = (lane < jb && c < jb) ? ap[(i64)c * lda] : T(0);
    });
    i64 bad = 0;
    static_for<0, 64>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k < jb) {                                   // uniform
            const T p = readlane(row[k], k);
            const bool ok = p != T(0) && isfinite(p);
            if (!ok && bad == 0) bad = j0 + k + 1;
            const T rinv = ok ? T(1) / p : T(0);
            const T l = row[k] * rinv;
            if (lane > k && ok) row[k] = l;
            static_for<k + 1, 64>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const T u = readlane(row[j], k);
                if (lane > k && ok) row[j] -= l * u;
            });
        }
    });
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        if (lane < jb && c < jb) ap[(i64)c * lda] = row[c];
    });
    if (lane < jb) piv[j0 + lane] = j0 + lane;
    if (lane == 0 && bad != 0)
        atomicMin(reinterpret_cast<unsigned long long*>(first_zero), (unsigned long long)bad);
}

// genp_l21_kernel: L21 := A21 U11^-1 for the rows [j0 + jb, nrows) below the
// diagonal block, one row per thread: U11 and the reciprocals of its
// diagonal in LDS (every lane reads the same word: broadcasts), the row in
// registers, column-oriented substitution.
template <typename T>
__global__ void __launch_bounds__(256)
genp_l21_kernel(T* __restrict__ A, i64 lda, i64 nrows, i64 j0, int jb)
{
    __shared__ __attribute__((aligned(16))) T s_u[64][64];     // s_u[p][k] = U11(p, k)
    __shared__ T s_r[64];
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        const int p = e & 63, k = e >> 6;
        s_u[p][k] = (p <= k && k < jb) ? A[(j0 + p) + (j0 + k) * lda] : T(0);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int p = threadIdx.x;
        const T d = s_u[p][p];
        s_r[p] = (p < jb && d != T(0) && isfinite(d)) ? T(1) / d : T(0);
    }
    __syncthreads();
    const i64 i = j0 + jb + (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    T* ap = A + i + j0 * lda;
    T x[64];
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        x[c] = c < jb ? ap[(i64)c * lda] : T(0);
    });
    int off = 0;
    static_for<0, 64>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        x[p] *= s_r[p];
        // row p of U11 is read once x[p] is final (an opaque dependency:
        // otherwise every read is hoisted to the top and the rows spill)
        __asm__ volatile("" : "+v"(off) : "v"(x[p]));
        const T* up = &s_u[0][0] + off + p * 64;
        static_for<p + 1, 64>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            x[k] -= x[p] * up[k];
        });
    });
    static_for<0, 64>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        if (c < jb) ap[(i64)c * lda] = x[c];
    });
}

__global__ void getrf_init_kernel(unsigned long long* tags, i64 ntags, int* err,
                                  i64* first_zero, i64 n, i64* perm)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntags) tags[i] = 0ull;
    if (i == 0) { *err = 0; *first_zero = n + 1; }
    if (perm != nullptr && i < n) perm[i] = i;
}

// info of the no-allocation factorisation, on the device: -1 when a bounded
// wait expired, else the 1-based first zero pivot or 0.
__global__ void getrf_finish_kernel(const int* err, const i64* first_zero, i64 n, i64* d_info)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *d_info = (err && *err) ? -1 : ((first_zero && *first_zero <= n) ? *first_zero : 0);
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

// Panel timer (diagnostics): while enabled, each panel-kernel launch of an
// allocating factorisation is bracketed by a pair of HIP events recorded on
// the stream the panel runs on, up to a fixed number of launches.
namespace {
std::mutex g_pt_mu;
std::vector<hipEvent_t> g_pt_ev;           // 2 per launch
size_t g_pt_used = 0;
bool g_pt_on = false;
}

static bool panel_timer_take(hipEvent_t* e0, hipEvent_t* e1)
{
    std::lock_guard<std::mutex> lk(g_pt_mu);
    if (!g_pt_on || 2 * g_pt_used + 1 >= g_pt_ev.size()) return false;
    *e0 = g_pt_ev[2 * g_pt_used];
    *e1 = g_pt_ev[2 * g_pt_used + 1];
    ++g_pt_used;
    return true;
}

// PA = LU of the m x n matrix A (rows [0, m), columns [0, n)): min(m, n)
// panel steps; the columns past min(m, n) of a wide matrix take the
// interchanges and the U solve, the rows past it of a tall one are only
// eliminated (LAPACK getrf and TMV's RecursiveLUDecompose semantics).
template <typename T, int NB>
int getrf_nb_impl(i64 m, i64 n, T* A, i64 lda, i64* piv, i64* perm, i64* info, hipStream_t s,
                  int lowprec, int pivoting, const GetrfNoAlloc* na)
{
    if (m < 0 || n < 0 || lda < (m > 0 ? m : 1) || (!info && !na) || (na && !na->d_info))
        return TMVX_ERR_ARG;
    if (info) *info = 0;
    const i64 R = m < n ? m : n;
    if (R > 0 && (!A || !piv)) return TMVX_ERR_ARG;
    if (perm && m != n) return TMVX_ERR_ARG;
    if (R == 0) {
        if (!na) return TMVX_OK;
        hipLaunchKernelGGL(getrf_finish_kernel, dim3(1), dim3(64), 0, s, nullptr, nullptr, R,
                           na->d_info);
        return hip_status(hipGetLastError());
    }
    constexpr int BB = PANEL_BB, NTH = PANEL_THREADS;
    typedef ColSlot<T, BB> CS;
    const int Gmax = (int)((m + NTH - 1) / NTH);
    if ((m + NTH - 1) / NTH > PANEL_GMAX) return TMVX_ERR_UNSUPPORTED;   // m > 262144

    const i64 ncslot = 2 * (i64)Gmax * CS::SIZE;
    const i64 nbslot = 2 * (i64)BB * NB * CS::WPV;
    const i64 nslot = ncslot + nbslot;
    const size_t szslots = sizeof(unsigned long long) * (size_t)nslot;
    const bool tslu = sizeof(T) == 4 && pivoting == PIVOT_TOURNAMENT && m == n;
    const bool nopiv = pivoting == PIVOT_NONE;
    if (nopiv && NB != 64) return TMVX_ERR_UNSUPPORTED;
    // one launch per panel for the interchanges + U12 inside an outer panel
    // and one per outer panel outside it (64-column panels); wider panels
    // take the per-panel interchange and U12 kernels
    const bool fused = NB == 64;
    constexpr int NBOM = sizeof(T) == 8 ? TMVX_NBO_MUL_F64 : TMVX_NBO_MUL;
    constexpr i64 NBO = NBOM * NB;
    static_assert(NBOM == 2 || NBOM == 4 || NBOM == 8, "panels per fused launch");
    const size_t szts = tslu ? sizeof(int) * (size_t)tslu_ws_ints(n) : 0;
    const size_t total = szslots + 256 + szts;
    char* ws = nullptr;
    if (na) {
        if (total > na->bytes || tslu || nopiv || lowprec != TMVX_LOWPREC_F32)
            return TMVX_ERR_UNSUPPORTED;
        ws = static_cast<char*>(na->ws);
    } else if (hipMallocAsync((void**)&ws, total, s) != hipSuccess) {
        return TMVX_ERR_ALLOC;
    }
    PanelWork wk;
    wk.cslots = reinterpret_cast<unsigned long long*>(ws);
    wk.bslots = wk.cslots + ncslot;
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
        const i64 cnt = R > nslot ? R : nslot;
        hipLaunchKernelGGL(getrf_init_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256),
                           0, s, wk.cslots, nslot, wk.err, wk.first_zero, R, perm);
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
        hipLaunchKernelGGL((laswp_panel_kernel<T, NB>), dim3(grid), dim3(256), 0, st, A, lda, m,
                           a0, a1, b0, b1, j0, jb, piv, pm, cpw);
        return hip_status(hipGetLastError());
    };
    // Interchanges of the panels in rows [Jp, Jp+rows) on columns [a0,a1) u
    // [b0,b1), then the U12 solve of those rows on [b0,b1) (fused path).
    // (solve_b = false: interchanges only, on both ranges)
    auto swap_solve = [&](i64 Jp, i64 rows, i64 a0, i64 a1, i64 b0, i64 b1, i64* pm,
                          hipStream_t st, bool solve_b = true) -> int {
        if (nopiv) {                       // no interchanges: only the solve on [b0, b1)
            if (!solve_b) return TMVX_OK;
            a1 = a0;
            pm = nullptr;
        }
        const i64 cols = (a1 - a0) + (b1 - b0);
        if (cols <= 0 && !pm) return TMVX_OK;
        // columns per wave: one (the interchanges are a chain of dependent
        // round trips per column: parallelism hides them), four for a solve
        // of many columns (its L rows are re-read per column)
        const int cpw = (solve_b && cols > 8192) ? 4 : 1;
        const i64 waves = (cols + cpw - 1) / cpw;
        const unsigned grid = (unsigned)(waves > 0 ? (waves + 3) / 4 : 1);
        hipLaunchKernelGGL((swap_solve_kernel<T, NBOM>), dim3(grid), dim3(256), 0, st, A, lda, m, Jp,
                           (int)rows, piv, a0, a1, b0, b1, cpw, pm, solve_b ? 1 : 0);
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
    SideSet* ss = nullptr;                 // taken from the pool on first need
    hipStream_t side = nullptr;
    hipEvent_t ev_main = nullptr, ev_side = nullptr;
    bool side_pending = false;
    for (i64 J0 = 0; J0 < R && rc == TMVX_OK; J0 += NBO) {
        const i64 JE = (R - J0) < NBO ? R : J0 + NBO;      // outer panel [J0, JE)
        for (i64 j0 = J0; j0 < JE && rc == TMVX_OK; j0 += NB) {
            const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
            if constexpr (sizeof(T) == 4) {
                if (tslu) {
                    // tournament pivots, their interchanges on the panel's
                    // own columns and the panel's L and U
                    if ((rc = tslu_panel<T>(A, lda, m, j0, jb, piv, wk.first_zero, ts_ws, s)) !=
                        TMVX_OK)
                        break;
                }
            }
            if (nopiv) {
                hipEvent_t pt0 = nullptr, pt1 = nullptr;
                const bool timed = panel_timer_take(&pt0, &pt1);
                if (timed) (void)hipEventRecord(pt0, s);
                hipLaunchKernelGGL((genp_diag_kernel<T>), dim3(1), dim3(64), 0, s, A, lda, j0, jb,
                                   piv, wk.first_zero);
                const i64 below = m - j0 - jb;
                if (below > 0)
                    hipLaunchKernelGGL((genp_l21_kernel<T>), dim3((unsigned)((below + 255) / 256)),
                                       dim3(256), 0, s, A, lda, m, j0, jb);
                if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
                if (timed) (void)hipEventRecord(pt1, s);
            } else if (!tslu) {
                const int G = (int)((m - j0 + NTH - 1) / NTH);
                hipEvent_t pt0 = nullptr, pt1 = nullptr;
                const bool timed = !na && panel_timer_take(&pt0, &pt1);
                if (timed) (void)hipEventRecord(pt0, s);
                if (G == 1)                    // every row in one workgroup: no exchange
                    hipLaunchKernelGGL((panel_kernel<T, NB, BB, NTH, true>), dim3(1), dim3(NTH),
                                       0, s, A, lda, m, j0, jb, piv, wk);
                else
                    hipLaunchKernelGGL((panel_kernel<T, NB, BB, NTH, false>), dim3(G), dim3(NTH),
                                       0, s, A, lda, m, j0, jb, piv, wk);
                if ((rc = hip_status(hipGetLastError())) != TMVX_OK) break;
                if (timed) (void)hipEventRecord(pt1, s);
            }
            // inside the outer panel only; the rest is deferred
            const i64 nr = JE - j0 - jb;
            if (fused) {
                if ((rc = swap_solve(j0, jb, J0, j0, j0 + jb, JE, perm, s)) != TMVX_OK) break;
            } else if ((rc = swap_cols(j0, jb, J0, j0, j0 + jb, JE, perm, s)) != TMVX_OK) {
                break;
            }
            if (nr > 0) {
                if (!fused && (rc = solve_rows(j0, jb, j0 + jb, nr, s)) != TMVX_OK) break;
                rc = gemm_launch<T>(0, 0, m - j0 - jb, nr, jb, T(-1), A + (j0 + jb) + j0 * lda,
                                    lda, A + j0 + (j0 + jb) * lda, lda, T(1),
                                    A + (j0 + jb) + (j0 + jb) * lda, lda, s);
            }
        }
        if (rc != TMVX_OK) break;
        if (side_pending) {
            if (side != s && hipStreamWaitEvent(s, ev_side, 0) != hipSuccess) {
                rc = TMVX_ERR_HIP;
                break;
            }
            side_pending = false;
        }
        // Deferred interchanges on the columns outside the outer panel.  With
        // a side stream to come, the caller's stream swaps only the next
        // outer panel's columns; the side stream swaps the rest (the left
        // L blocks and the far trailing columns) before its own work.
        const i64 nr = n - JE;                              // columns right of the panel
        const i64 nrow = m - JE;                            // rows below it
        const i64 JE2 = (n - JE) < NBO ? n : JE + NBO;      // next outer panel [JE, JE2)
        const bool has_side = nr > 0 && JE2 < n;
        if (fused && TMVX_K2_FUSED) {
            // one launch: the outer panel's interchanges and its U12 block
            // row on the next outer panel's columns (and, without a side
            // stream to come, the interchanges on the left blocks)
            rc = swap_solve(J0, JE - J0, 0, has_side ? 0 : J0, JE, JE2, nullptr, s);
            if (nr <= 0 || rc != TMVX_OK) continue;
        } else if (fused) {
            // one launch for the interchanges, then the block row's U12 on
            // the trsm / GEMM kernels
            rc = swap_solve(J0, JE - J0, 0, has_side ? 0 : J0, JE, JE2, nullptr, s, false);
            if (nr <= 0 || rc != TMVX_OK) continue;
            if ((rc = block_row_u12(J0, JE, JE, JE2, s)) != TMVX_OK) break;
        } else {
            for (i64 j0 = J0; j0 < JE && rc == TMVX_OK; j0 += NB) {
                const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
                rc = has_side ? swap_cols(j0, jb, JE, JE2, JE2, JE2, nullptr, s)
                              : swap_cols(j0, jb, 0, J0, JE, n, nullptr, s);
            }
            if (nr <= 0 || rc != TMVX_OK) continue;
            // next panel's columns on the caller's stream
            if ((rc = block_row_u12(J0, JE, JE, JE2, s)) != TMVX_OK) break;
        }
        if (nrow > 0)
            rc = trailing(nrow, JE2 - JE, JE - J0, A + JE + J0 * lda, A + J0 + JE * lda,
                          A + JE + JE * lda, bmax_main, s);
        if (rc != TMVX_OK || JE2 >= n) continue;
        // the rest of the trailing matrix on the side stream (without
        // allocation: the same kernels in the same order on s itself)
        if (!side) {
            if (na) {
                side = s;
            } else {
                if ((rc = side_acquire(&ss)) != TMVX_OK) break;
                side = ss->stream;
                ev_main = ss->ev_main;
                ev_side = ss->ev_side;
            }
        }
        if (side != s && (hipEventRecord(ev_main, s) != hipSuccess ||
                          hipStreamWaitEvent(side, ev_main, 0) != hipSuccess)) {
            rc = TMVX_ERR_HIP;
            break;
        }
        if (fused) {
            // many columns: one launch for the interchanges, the U12 block
            // row on the trsm (L rows kept in registers across 16 columns)
            // and GEMM kernels, which move L once per 16 columns instead of
            // once per column
            if ((rc = swap_solve(J0, JE - J0, 0, J0, JE2, n, nullptr, side, false)) != TMVX_OK)
                break;
            if ((rc = block_row_u12(J0, JE, JE2, n, side)) != TMVX_OK) break;
        } else {
            for (i64 j0 = J0; j0 < JE && rc == TMVX_OK; j0 += NB) {
                const int jb = (int)((JE - j0) < NB ? (JE - j0) : NB);
                rc = swap_cols(j0, jb, 0, J0, JE2, n, nullptr, side);
            }
            if (rc != TMVX_OK) break;
            if ((rc = block_row_u12(J0, JE, JE2, n, side)) != TMVX_OK) break;
        }
        if (nrow > 0)
            rc = trailing(nrow, n - JE2, JE - J0, A + JE + J0 * lda, A + J0 + JE2 * lda,
                          A + JE + JE2 * lda, bmax_side, side);
        if (rc != TMVX_OK) break;
        if (side != s && hipEventRecord(ev_side, side) != hipSuccess) { rc = TMVX_ERR_HIP; break; }
        side_pending = true;
    }
    if (na) {                              // asynchronous: info stays on the device
        if (rc != TMVX_OK) return rc;
        hipLaunchKernelGGL(getrf_finish_kernel, dim3(1), dim3(64), 0, s, wk.err, wk.first_zero, R,
                           na->d_info);
        return hip_status(hipGetLastError());
    }
    // whatever happened above, the caller's stream waits for the side stream
    if (side) {
        if (hipEventRecord(ev_side, side) != hipSuccess ||
            hipStreamWaitEvent(s, ev_side, 0) != hipSuccess)
            rc = rc == TMVX_OK ? TMVX_ERR_HIP : rc;
    }
    int herr = 0;
    i64 fz = R + 1;
    if (hipMemcpyAsync(&herr, wk.err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&fz, wk.first_zero, sizeof(i64), hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = TMVX_ERR_HIP;
    (void)hipFreeAsync(ws, s);
    const bool sync_ok = hipStreamSynchronize(s) == hipSuccess;
    side_release(ss);                      // its work is done: s waited for it
    if (!sync_ok) return TMVX_ERR_HIP;
    if (rc != TMVX_OK) return rc;
    if (herr) return TMVX_ERR_TIMEOUT;
    *info = fz <= R ? fz : 0;
    return TMVX_OK;
}

template <typename T>
int getrf_mn_impl(i64 m, i64 n, T* A, i64 lda, i64* piv, i64* perm, i64* info, hipStream_t s,
                  int lowprec, int pivoting, const GetrfNoAlloc* na)
{
    // tournament and no-interchange panels are 64 columns wide
    if constexpr (sizeof(T) == 4 && TMVX_F32_NB != PANEL_NB) {
        if (pivoting == PIVOT_PARTIAL)
            return getrf_nb_impl<T, TMVX_F32_NB>(m, n, A, lda, piv, perm, info, s, lowprec,
                                                 pivoting, na);
    }
    return getrf_nb_impl<T, PANEL_NB>(m, n, A, lda, piv, perm, info, s, lowprec, pivoting, na);
}

template <typename T>
int getrf_impl(i64 n, T* A, i64 lda, i64* piv, i64* perm, i64* info, hipStream_t s, int lowprec,
               int pivoting, const GetrfNoAlloc* na)
{
    return getrf_mn_impl<T>(n, n, A, lda, piv, perm, info, s, lowprec, pivoting, na);
}

// Interchange list -> composed permutation, without a chain of swaps.
// Step k swaps positions k and p_k >= k, so position k is final after step
// k, and a position q > k is touched before step k only by the steps j < k
// with p_j = q (the sources of q).  With f(q) = the content of position q
// just before step q:
//   f(q) = f(last source of q), or q when q has no source;
//   perm[k] = f(k) when p_k = k, else the content of p_k before step k:
//   f(previous source of p_k), or p_k when k is its first source.
// One workgroup sorts the keys (p_k, k) of the sources in LDS (bitonic),
// reads every source's predecessor and every target's last source off the
// sorted runs, resolves f by pointer jumping (parents point to smaller
// indices, so ceil(log2 n) rounds suffice; it stops when a round changes
// nothing) and gathers.  The 8n bytes of perm serve as the workspace: int
// 2q holds q's parent, int 2k + 1 what perm[k] takes (an index into f, or
// -(p_k + 1) for the literal p_k), and each thread reads its own pair before
// it writes perm[k] over it.  n > PERM_LDS_MAX, or a list with some
// 0 <= piv[k] < k (not an LU interchange list): one thread, in order.
constexpr int PERM_LDS_MAX = 32768;
constexpr int PERM_KBITS = 15;                 // k < 2^15 in a key
static_assert(PERM_LDS_MAX <= (1 << PERM_KBITS), "key packing");

__global__ void __launch_bounds__(1024)
piv_to_perm_kernel(i64 n, const i64* __restrict__ piv, i64* __restrict__ perm)
{
    __shared__ int sk[PERM_LDS_MAX];
    __shared__ int s_back;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const bool lds = n <= PERM_LDS_MAX;
    if (tid == 0) s_back = 0;
    __syncthreads();
    int* ws = reinterpret_cast<int*>(perm);
    const int n32 = lds ? (int)n : 0;
    int N2 = 1;
    while (N2 < n32) N2 <<= 1;
    if (lds) {
        for (int k = tid; k < N2; k += nt) {
            int key = INT_MAX;                     // not a source: sorts last
            if (k < n32) {
                const i64 p = piv[k];
                if (p >= 0 && p < k) s_back = 1;
                if (p > k && p < n) key = ((int)p << PERM_KBITS) | k;
                ws[2 * k] = k;                     // no source yet: f(k) = k
                ws[2 * k + 1] = k;                 // p_k = k: perm[k] = f(k)
            }
            sk[k] = key;
        }
    } else {
        for (i64 i = tid; i < n; i += nt) {
            const i64 p = piv[i];
            if (p >= 0 && p < i) s_back = 1;
        }
    }
    __syncthreads();
    if (!lds || s_back) {
        for (i64 i = tid; i < n; i += nt) {
            if (lds) sk[i] = (int)i; else perm[i] = i;
        }
        __syncthreads();
        if (tid == 0) {
            for (i64 k = 0; k < n; ++k) {
                const i64 p = piv[k];
                if (p == k || p < 0 || p >= n) continue;
                if (lds) { const int x = sk[k]; sk[k] = sk[p]; sk[p] = x; }
                else { const i64 x = perm[k]; perm[k] = perm[p]; perm[p] = x; }
            }
        }
        __syncthreads();
        if (lds)
            for (i64 i = tid; i < n; i += nt) perm[i] = sk[i];
        return;
    }
    // bitonic sort of the N2 keys, ascending
    for (int size = 2; size <= N2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int idx = tid; idx < (N2 >> 1); idx += nt) {
                const int i = 2 * idx - (idx & (stride - 1));
                const int a = sk[i], b = sk[i + stride];
                if ((a > b) == ((i & size) == 0)) { sk[i] = b; sk[i + stride] = a; }
            }
            __syncthreads();
        }
    }
    // runs of one target: predecessor of each source, last source of each target
    constexpr int KMASK = (1 << PERM_KBITS) - 1;
    for (int i = tid; i < N2; i += nt) {
        const int key = sk[i];
        if (key == INT_MAX) continue;
        const int t = key >> PERM_KBITS, k = key & KMASK;
        const int kp = i > 0 ? sk[i - 1] : INT_MAX;
        const int kn = i + 1 < N2 ? sk[i + 1] : INT_MAX;
        ws[2 * k + 1] = (kp != INT_MAX && (kp >> PERM_KBITS) == t) ? (kp & KMASK) : -(t + 1);
        if (kn == INT_MAX || (kn >> PERM_KBITS) != t) ws[2 * t] = k;
    }
    __syncthreads();
    for (int q = tid; q < n32; q += nt) sk[q] = ws[2 * q];
    __syncthreads();
    for (int r = 0; (1 << r) < n32; ++r) {     // pointer jumping to the roots
        int changed = 0;
        for (int q = tid; q < n32; q += nt) {
            const int a = sk[q];
            const int b = sk[a];
            if (b != a) { sk[q] = b; changed = 1; }
        }
        if (!__syncthreads_or(changed)) break;
    }
    for (int k = tid; k < n32; k += nt) {
        const int c = ws[2 * k + 1];
        perm[k] = c >= 0 ? (i64)sk[c] : (i64)(-(c + 1));
    }
}

int piv_to_perm(i64 n, const i64* piv, i64* perm, hipStream_t s)
{
    if (n <= 0) return TMVX_OK;
    hipLaunchKernelGGL(piv_to_perm_kernel, dim3(1), dim3(1024), 0, s, n, piv, perm);
    return hip_status(hipGetLastError());
}

template int getrf_mn_impl<double>(i64, i64, double*, i64, i64*, i64*, i64*, hipStream_t, int,
                                   int, const GetrfNoAlloc*);
template int getrf_mn_impl<float>(i64, i64, float*, i64, i64*, i64*, i64*, hipStream_t, int, int,
                                  const GetrfNoAlloc*);
template int getrf_impl<double>(i64, double*, i64, i64*, i64*, i64*, hipStream_t, int, int,
                                const GetrfNoAlloc*);
template int getrf_impl<float>(i64, float*, i64, i64*, i64*, i64*, hipStream_t, int, int,
                               const GetrfNoAlloc*);

// Scratch of the no-allocation factorisation: the panel exchange slots and
// flags in a static device array (each GPU holds its own copy for the
// life of the process).  Host code never uses the array's name as an
// address: the device address comes from hipGetSymbolAddress, once per
// device, with that device current.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64 * 1024;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};
}

static int scratch_address(void** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TMVX_ERR_HIP;
    if (dev < 0 || dev >= 64) return TMVX_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TMVX_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TMVX_OK;
}

} // namespace tmvx

extern "C" int tmvx_panel_timer_start(int cap)
{
    using namespace tmvx;
    if (cap <= 0 || cap > (1 << 20)) return TMVX_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_pt_mu);
    if (g_pt_on) return TMVX_ERR_ARG;
    g_pt_ev.assign(2 * (size_t)cap, nullptr);
    for (auto& e : g_pt_ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            for (auto& f : g_pt_ev)
                if (f) (void)hipEventDestroy(f);
            g_pt_ev.clear();
            return TMVX_ERR_HIP;
        }
    }
    g_pt_used = 0;
    g_pt_on = true;
    return TMVX_OK;
}

extern "C" int tmvx_panel_timer_stop(int64_t* launches, double* total_ms)
{
    using namespace tmvx;
    std::lock_guard<std::mutex> lk(g_pt_mu);
    if (!g_pt_on || !launches || !total_ms) return TMVX_ERR_ARG;
    int rc = TMVX_OK;
    double sum = 0.0;
    for (size_t i = 0; i < g_pt_used && rc == TMVX_OK; ++i) {
        float ms = 0.0f;
        if (hipEventSynchronize(g_pt_ev[2 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, g_pt_ev[2 * i], g_pt_ev[2 * i + 1]) != hipSuccess)
            rc = TMVX_ERR_HIP;
        sum += ms;
    }
    *launches = (int64_t)g_pt_used;
    *total_ms = sum;
    for (auto& e : g_pt_ev) (void)hipEventDestroy(e);
    g_pt_ev.clear();
    g_pt_used = 0;
    g_pt_on = false;
    return rc;
}

#ifdef TMVX_PANEL_TRACE
extern "C" int tmvx_panel_trace_set(unsigned long long* buf, long long j0)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(tmvx::g_panel_trace), &buf, sizeof(buf)) == hipSuccess &&
                   hipMemcpyToSymbol(HIP_SYMBOL(tmvx::g_panel_trace_j0), &j0, sizeof(j0)) ==
                       hipSuccess
               ? 0 : -3;
}
#endif

extern "C" {

int tmvx_getrf_f64(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return tmvx::getrf_impl<double>(n, A, lda, piv, nullptr, info, (hipStream_t)stream); }
int tmvx_getrf_f32(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info, void* stream)
{ return tmvx::getrf_impl<float>(n, A, lda, piv, nullptr, info, (hipStream_t)stream); }

int tmvx_getrf_f64_noalloc(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* d_info,
                           void* stream)
{
    if (n > 65536) return TMVX_ERR_UNSUPPORTED;              // the scratch array's limit
    if (n < 0 || lda < (n > 0 ? n : 1) || !d_info || (n > 0 && (!A || !piv)))
        return TMVX_ERR_ARG;
    void* ws = nullptr;
    const int rc = tmvx::scratch_address(&ws);
    if (rc != TMVX_OK) return rc;
    const tmvx::GetrfNoAlloc na{ws, tmvx::NOALLOC_SCRATCH_BYTES, d_info};
    return tmvx::getrf_impl<double>(n, A, lda, piv, nullptr, nullptr, (hipStream_t)stream,
                                    TMVX_LOWPREC_F32, 0, &na);
}

int tmvx_getrf_mn_f64(int64_t m, int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info,
                      void* stream)
{
    return tmvx::getrf_mn_impl<double>(m, n, A, lda, piv, nullptr, info, (hipStream_t)stream,
                                       TMVX_LOWPREC_F32, 0, nullptr);
}

int tmvx_getrf_mn_f32(int64_t m, int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info,
                      void* stream)
{
    return tmvx::getrf_mn_impl<float>(m, n, A, lda, piv, nullptr, info, (hipStream_t)stream,
                                      TMVX_LOWPREC_F32, 0, nullptr);
}

int tmvx_getrf_f32_tournament(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info,
                              void* stream)
{
    return tmvx::getrf_impl<float>(n, A, lda, piv, nullptr, info, (hipStream_t)stream,
                                   TMVX_LOWPREC_F32, tmvx::PIVOT_TOURNAMENT);
}

int tmvx_getrf_nopiv_f64(int64_t n, double* A, int64_t lda, int64_t* piv, int64_t* info,
                         void* stream)
{
    return tmvx::getrf_impl<double>(n, A, lda, piv, nullptr, info, (hipStream_t)stream,
                                    TMVX_LOWPREC_F32, tmvx::PIVOT_NONE);
}

int tmvx_getrf_nopiv_f32(int64_t n, float* A, int64_t lda, int64_t* piv, int64_t* info,
                         void* stream)
{
    return tmvx::getrf_impl<float>(n, A, lda, piv, nullptr, info, (hipStream_t)stream,
                                   TMVX_LOWPREC_F32, tmvx::PIVOT_NONE);
}

} // extern "C"
