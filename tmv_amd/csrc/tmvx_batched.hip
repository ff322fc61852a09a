// Batched LU factorisation and solve of many small independent systems, and
// a batched GEMM over the same kind of ragged sizes.
//
// Counterpart, per system, of TMV's LU_Decompose + LU_LDivEq
// (src/TMV_LUDecompose.cpp:217-407, src/TMV_LUDiv.cpp:185-226) and MultMM
// (src/TMV_MultMM.cpp:404-477); TMV has no batched entry point, a caller
// loops over systems.  Here one persistent launch walks a work list.
//
// Factor + solve kernel (one system per workgroup or per wave):
//   The system's matrix lives in registers, distributed 2D-cyclically over
//   a GD x GD thread grid: thread (rb, cb) owns rows rb + GD p and columns
//   cb + GD q (p, q < TR), so the shrinking trailing matrix stays spread
//   over all threads.  The right-hand sides are one more column level
//   (thread cb holds RHS column cb), so forward elimination of B happens
//   with the factorisation.  The column levels rotate through register
//   slot 0 (one runtime loop body, not one per level), and the pivot row is
//   read by a uniform register index.  Rows are never moved: pivoting is implicit,
//   each wave keeps the position <-> row maps in registers and applies the
//   interchange list's semantics (TMV's / LAPACK's piv[j]), so pivots, tie
//   rule (first maximum in the current row order) and TMV's underflow rule
//   are those of TMV's GEPP.  Per column: the owners of column j publish it
//   to LDS; every wave finds the pivot by DPP argmax; the pivot row and the
//   multipliers (reciprocal of the pivot times the column, as TMV scales)
//   go to LDS; every thread updates its tile.  Two barriers per column.
//   The factors are written out in pivoted order; back substitution with U
//   runs one wave per right-hand side, lane per row, from the just-written
//   factors.
//   Two schedules share one launch: systems with n <= split run one per
//   wave (a GDw = 8 grid, wave-local synchronisation), larger ones one per
//   workgroup (GDg = 16, 256 threads).  A plan (made once per batch shape)
//   orders the work list largest first and picks the schedule pair, the
//   split and the grid, by timing the candidates when asked to.
// The launch allocates nothing and sets nothing on the host, so it can be
// captured in a hipGraph and replayed.
#include "tmvx_lu.h"
#include "tmvx_pivot.h"
#include "tmvx_wave_solve.h"
#include "../../include/tmvx_batched.h"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

namespace tmvx {

template <int GD_, int TR_>
struct BSched {
    static constexpr int GD = GD_, TR = TR_, NT = GD_ * GD_, NMAX = GD_ * TR_;
    static constexpr int PR = (NMAX + 63) / 64;      // map entries per lane
    static constexpr int NW = (NT + 63) / 64;        // waves per system
};

template <typename T, int NMAX, int GD>
struct BLds {
    T col[NMAX];            // column j, indexed by row
    T urow[NMAX + GD];      // pivot row: columns, then right-hand sides
    T mult[NMAX];           // multipliers, indexed by row
    int pos[NMAX];          // final position of each row
    int piv[NMAX];          // interchange list, written out at the end
};

struct GroupSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};
struct WaveSync {
    __device__ __forceinline__ void operator()() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

template <typename T>
struct BatchArgs {
    const int* order;       // system index per work-list slot
    const int4* items;      // {mode (0 group, 1 wave), first slot, count, 0}
    int nitems;
    const int* n;
    const int* nrhs;
    T* const* A;            // n x n column-major, lda = n
    i64* const* piv;        // n interchanges, 0-based
    T* const* B;            // n x nrhs column-major, ldb = n (may be null if nrhs = 0)
    int* info;              // per system: 1-based first zero pivot, or 0
};

// Value of slot s = sel (uniform) among a lane's PR map entries.
template <int PR>
__device__ __forceinline__ int pick(const int (&v)[PR], int sel)
{
    int r = v[0];
    static_for<1, PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if (sel == s) r = v[s];
    });
    return r;
}

// The TR rows a thread holds in one column level, as a vector so that the
// pivot row (a uniform slot chosen at run time) is read with register-
// indexed moves instead of a chain of selects.
template <typename T, int TR> struct RowVec {
    typedef T type __attribute__((ext_vector_type(TR)));
};

// Factor (and solve, nr > 0) one system with schedule S; tid in [0, S::NT).
template <typename T, class S, class Sync>
__device__ __forceinline__ void lu_solve_one(int n, int nr, T* __restrict__ A, i64* __restrict__ piv,
                             T* __restrict__ B, int* __restrict__ info,
                             BLds<T, S::NMAX, S::GD>& L, int tid, Sync sync)
{
    constexpr int GD = S::GD, TR = S::TR, PR = S::PR, NW = S::NW;
    typedef typename RowVec<T, TR>::type V;
    const int rb = tid / GD, cb = tid % GD;
    const int lane = tid & 63, wave = tid >> 6;
    const i64 ld = n;

    // a[s][p] = row rb + GD p of column cb + GD ((s + qj) mod TR): while
    // column level qj is eliminated it sits in slot 0, the levels still to
    // come in slots 1 .. TR-1-qj; the slots rotate once per level, so after
    // TR levels every column is back in its own slot.
    V a[TR];
    V y;
    unsigned act = 0;        // bit p: row rb + GD p not yet pivotal
    static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
        constexpr int p = decltype(pc)::value;
        const int r = rb + GD * p;
        static_for<0, TR>([&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            const int c = cb + GD * q;
            a[q][p] = (r < n && c < n) ? A[r + c * ld] : T(0);
        });
        y[p] = (r < n && cb < nr) ? B[r + cb * ld] : T(0);
        if (r < n) act |= 1u << p;
    });
    // Position <-> row maps, a full copy in every wave: lane l holds the
    // entries l + 64 s.  phys[k] = row at position k, pos[r] = position of r.
    int pos_r[PR], phys_r[PR];
    static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        pos_r[s] = phys_r[s] = lane + 64 * s;
    });
    int first_zero = 0;

    for (int qj = 0; qj < TR; ++qj) {
        for (int jj = 0; jj < GD; ++jj) {
            const int j = qj * GD + jj;
            if (j >= n) break;
            // 1. column j (slot 0) to LDS
            if (cb == jj) {
                static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
                    constexpr int p = decltype(pc)::value;
                    const int r = rb + GD * p;
                    if (r < n) L.col[r] = a[0][p];
                });
            }
            sync();
            // 2. pivot: largest |a| over rows at positions >= j, first in
            //    position order on ties (every wave computes the same)
            T bv = T(-1);
            int bi = INT_MAX;
            static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                const int r = lane + 64 * s;
                if (r < n && pos_r[s] >= j) {
                    const T v = fabs(L.col[r]);
                    if (better(v, pos_r[s], bv, bi)) { bv = v; bi = pos_r[s]; }
                }
            });
            wave_argmax(bv, bi);
            const bool zero = !(bv >= tiny<T>()) || bi == INT_MAX;
            const int lp = zero ? j : bi;
            const int rp = __builtin_amdgcn_readlane(pick(phys_r, lp >> 6), lp & 63);
            const int rj = __builtin_amdgcn_readlane(pick(phys_r, j >> 6), j & 63);
            static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                const int k = lane + 64 * s;
                if (k == j) phys_r[s] = rp; else if (k == lp) phys_r[s] = rj;
                if (k == rp) pos_r[s] = j; else if (k == rj) pos_r[s] = lp;
            });
            const T rcp = zero ? T(0) : T(1) / L.col[rp];
            // (kept in LDS: a global store here would make every barrier
            // wait for its completion)
            if (tid == 0) L.piv[j] = lp;
            if (zero && first_zero == 0) first_zero = j + 1;
            // 3. pivot row (register slot pp, read by index) and multipliers
            const int pp = rp / GD;
            const bool own_piv = rb == rp % GD;
            if (own_piv) {
                act &= ~(1u << pp);
                static_for<0, TR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    if (s < TR - qj) {
                        const int c = cb + GD * (s + qj);
                        if (c > j && c < n) L.urow[c] = a[s][pp];
                    }
                });
                if (cb < nr) L.urow[n + cb] = y[pp];
            }
            if (cb == jj) {
                static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
                    constexpr int p = decltype(pc)::value;
                    const bool isact = (act >> p) & 1u;
                    const bool zp = zero && own_piv && p == pp;   // zeroed diagonal
                    const T m = zero ? T(0) : a[0][p] * rcp;
                    a[0][p] = (isact || zp) ? m : a[0][p];
                    if (isact) L.mult[rb + GD * p] = m;
                });
            }
            sync();
            // 4. rank-1 update, unpredicated: pivoted rows get multiplier 0
            //    and finished columns pivot-row value 0, so their FMAs add
            //    exactly -0 (finite data; non-finite inputs give non-finite
            //    factors, as in any LU).  Skipped for a zero column.
            if (!zero) {
                T l[TR];
                static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
                    constexpr int p = decltype(pc)::value;
                    l[p] = ((act >> p) & 1u) ? L.mult[rb + GD * p] : T(0);
                });
                static_for<0, TR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    if (s < TR - qj) {
                        const int c = cb + GD * (s + qj);
                        const T u = (c > j && c < n) ? L.urow[c] : T(0);
                        static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
                            constexpr int p = decltype(pc)::value;
                            a[s][p] = fma(-l[p], u, a[s][p]);
                        });
                    }
                });
                const T uy = cb < nr ? L.urow[n + cb] : T(0);
                static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
                    constexpr int p = decltype(pc)::value;
                    y[p] = fma(-l[p], uy, y[p]);
                });
            }
        }
        // next column level into slot 0
        const V t0 = a[0];
        static_for<0, TR - 1>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            a[s] = a[s + 1];
        });
        a[TR - 1] = t0;
    }

    // Write the factors (and L^-1 P B) in pivoted row order.
    if (tid < 64) {
        static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            const int r = lane + 64 * s;
            if (r < n) L.pos[r] = pos_r[s];
        });
    }
    sync();
    static_for<0, TR>([&](auto pc) __attribute__((always_inline)) {
        constexpr int p = decltype(pc)::value;
        const int r = rb + GD * p;
        if (r < n) {
            const i64 lp = L.pos[r];
            static_for<0, TR>([&](auto qc) __attribute__((always_inline)) {
                constexpr int q = decltype(qc)::value;
                const int c = cb + GD * q;
                if (c < n) A[lp + c * ld] = a[q][p];
            });
            if (cb < nr) B[lp + cb * ld] = y[p];
        }
    });
    for (int k = tid; k < n; k += S::NT) piv[k] = L.piv[k];
    if (tid == 0) *info = first_zero;
    sync();
    if (first_zero != 0) return;        // no solution for a singular system

    // Back substitution U x = y, one wave per right-hand side.
    for (int c = wave; c < nr; c += NW) {
        T* Bc = B + (i64)c * ld;
        T yv[PR], rc[PR];
        static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            const int i = lane + 64 * s;
            yv[s] = i < n ? Bc[i] : T(0);
            rc[s] = i < n ? T(1) / A[i + i * ld] : T(0);
        });
        constexpr int CH = 8;
        for (int k0 = n - 1; k0 >= 0; k0 -= CH) {
            T colv[CH][PR];
            static_for<0, CH>([&](auto uc) __attribute__((always_inline)) {
                constexpr int u = decltype(uc)::value;
                const int k = k0 - u;
                static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    const int i = lane + 64 * s;
                    colv[u][s] = (k >= 0 && i < k) ? A[i + (i64)k * ld] : T(0);
                });
            });
            static_for<0, CH>([&](auto uc) __attribute__((always_inline)) {
                constexpr int u = decltype(uc)::value;
                const int k = k0 - u;
                if (k >= 0) {
                    T ysel = yv[0], rsel = rc[0];
                    static_for<1, PR>([&](auto sc) __attribute__((always_inline)) {
                        constexpr int s = decltype(sc)::value;
                        if ((k >> 6) == s) { ysel = yv[s]; rsel = rc[s]; }
                    });
                    const T yk = readlane(ysel, k & 63);
                    const T xk = yk == T(0) ? T(0) : yk * readlane(rsel, k & 63);
                    static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
                        constexpr int s = decltype(sc)::value;
                        const int i = lane + 64 * s;
                        if (i == k) yv[s] = xk;
                        else yv[s] -= colv[u][s] * xk;
                    });
                }
            });
        }
        static_for<0, PR>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            const int i = lane + 64 * s;
            if (i < n) Bc[i] = yv[s];
        });
    }
}

template <typename T, class SW, class SG>
struct BatchKernelLds {
    static constexpr int NWAVE = SG::NT / 64;
    union {
        BLds<T, SG::NMAX, SG::GD> g;
        BLds<T, SW::NMAX, SW::GD> w[NWAVE];
    };
};

template <typename T, class SW, class SG>
__global__ void __launch_bounds__(SG::NT)
batched_lu_solve_kernel(BatchArgs<T> args)
{
    static_assert(SW::NT == 64, "wave schedule: one wave per system");
    __shared__ BatchKernelLds<T, SW, SG> lds;
    for (int it = blockIdx.x; it < args.nitems; it += gridDim.x) {
        const int4 item = args.items[it];
        if (item.x == 0) {
            const int idx = args.order[item.y];
            lu_solve_one<T, SG>(args.n[idx], args.nrhs[idx], args.A[idx], args.piv[idx],
                                args.B ? args.B[idx] : nullptr, args.info + idx, lds.g,
                                (int)threadIdx.x, GroupSync());
        } else {
            const int w = (int)threadIdx.x >> 6;
            if (w < item.z) {
                const int idx = args.order[item.y + w];
                lu_solve_one<T, SW>(args.n[idx], args.nrhs[idx], args.A[idx], args.piv[idx],
                                    args.B ? args.B[idx] : nullptr, args.info + idx, lds.w[w],
                                    (int)threadIdx.x & 63, WaveSync());
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Batched solve with stored factors: one wave per (system, right-hand side).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
batched_getrs_kernel(int trans, int count, const int* __restrict__ nv,
                     const int* __restrict__ nrhsv, const T* const* __restrict__ LUv,
                     const i64* const* __restrict__ pivv, T* const* __restrict__ Bv)
{
    const int lane = threadIdx.x & 63;
    const int gw = (int)((blockIdx.x * (i64)blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * (i64)blockDim.x) >> 6);
    for (int sys = gw; sys < count; sys += nwaves) {
        const int n = nv[sys], nr = nrhsv[sys];
        if (n < 0 || n > WS_NMAX) continue;
        for (int c = 0; c < nr; ++c) {
            T* b = Bv[sys] + (i64)c * n;
            T x[WS_PR];
            wave_load(x, b, n, lane);
            wave_getrs(trans, n, LUv[sys], (i64)n, pivv[sys], x, lane);
            wave_store(x, b, n, lane);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched GEMM C_i = alpha A_i B_i + beta C_i (m_i x k_i times k_i x n_i,
// column-major, packed: lda = m, ldb = k, ldc = m), one workgroup per
// problem, each wave a 32 x 32 block of C on MFMA.  The k index inside an
// MFMA step is permuted (k = k0 + 4g + s) so a lane reads 4 consecutive k
// of its B column.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
batched_gemm_kernel(int count, const int* __restrict__ mv, const int* __restrict__ nv,
                    const int* __restrict__ kv, T alpha, const T* const* __restrict__ Av,
                    const T* const* __restrict__ Bv, T beta, T* const* __restrict__ Cv)
{
    typedef typename Mfma<T>::v4 v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    for (int idx = blockIdx.x; idx < count; idx += gridDim.x) {
        const int M = mv[idx], N = nv[idx], K = kv[idx];
        const T* A = Av[idx];
        const T* B = Bv[idx];
        T* C = Cv[idx];
        const int nbm = (M + 31) / 32, nbn = (N + 31) / 32;
        for (int blk = wave; blk < nbm * nbn; blk += 4) {
            const int r0 = (blk % nbm) * 32, c0 = (blk / nbm) * 32;
            v4 acc[2][2];
            static_for<0, 2>([&](auto ic) __attribute__((always_inline)) {
                static_for<0, 2>([&](auto jc) __attribute__((always_inline)) {
                    acc[decltype(ic)::value][decltype(jc)::value] = v4{0, 0, 0, 0};
                });
            });
            for (int k0 = 0; k0 < K; k0 += 16) {
                T af[2][4], bf[2][4];
                static_for<0, 4>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    const int k = k0 + 4 * g + s;
                    static_for<0, 2>([&](auto tc) __attribute__((always_inline)) {
                        constexpr int t = decltype(tc)::value;
                        const int r = r0 + 16 * t + li;
                        const int c = c0 + 16 * t + li;
                        af[t][s] = (k < K && r < M) ? A[r + (i64)k * M] : T(0);
                        bf[t][s] = (k < K && c < N) ? B[k + (i64)c * K] : T(0);
                    });
                });
                static_for<0, 4>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    static_for<0, 2>([&](auto ic) __attribute__((always_inline)) {
                        constexpr int i = decltype(ic)::value;
                        static_for<0, 2>([&](auto jc) __attribute__((always_inline)) {
                            constexpr int j = decltype(jc)::value;
                            acc[i][j] = Mfma<T>::mac(af[i][s], bf[j][s], acc[i][j]);
                        });
                    });
                });
            }
            static_for<0, 2>([&](auto ic) __attribute__((always_inline)) {
                constexpr int i = decltype(ic)::value;
                static_for<0, 2>([&](auto jc) __attribute__((always_inline)) {
                    constexpr int j = decltype(jc)::value;
                    const int c = c0 + 16 * j + li;
                    static_for<0, 4>([&](auto rc) __attribute__((always_inline)) {
                        constexpr int r = decltype(rc)::value;
                        const int row = r0 + 16 * i + Mfma<T>::row(g, r);
                        if (row < M && c < N) {
                            T* p = C + row + (i64)c * M;
                            const T v = alpha * acc[i][j][r];
                            *p = beta == T(0) ? v : v + beta * *p;
                        }
                    });
                });
            });
        }
    }
}

// ---------------------------------------------------------------------------
// Host side: schedule table, plans, autotuning.
// ---------------------------------------------------------------------------
typedef BSched<16, 8> SG128;           // 256 threads, n <= 128
typedef BSched<8, 2> SW16;             // one wave, n <= 16
typedef BSched<8, 4> SW32;             // one wave, n <= 32
typedef BSched<8, 8> SW64;             // one wave, n <= 64

struct BatchKernelInfo {
    const char* name;
    int wave_nmax, group_nmax, group_threads, wave_rhsmax, group_rhsmax;
    const void* f64;
    const void* f32;
};

static const BatchKernelInfo kBatchKernels[] = {
    {"wave n<=16 + group 16x16 threads n<=128", 16, 128, 256, 8, 16,
     (const void*)&batched_lu_solve_kernel<double, SW16, SG128>,
     (const void*)&batched_lu_solve_kernel<float, SW16, SG128>},
    {"wave n<=32 + group 16x16 threads n<=128", 32, 128, 256, 8, 16,
     (const void*)&batched_lu_solve_kernel<double, SW32, SG128>,
     (const void*)&batched_lu_solve_kernel<float, SW32, SG128>},
    {"wave n<=64 + group 16x16 threads n<=128", 64, 128, 256, 8, 16,
     (const void*)&batched_lu_solve_kernel<double, SW64, SG128>,
     (const void*)&batched_lu_solve_kernel<float, SW64, SG128>},
};
constexpr int kNumBatchKernels = sizeof(kBatchKernels) / sizeof(kBatchKernels[0]);

} // namespace tmvx

struct tmvx_batch_plan_s {
    int dtype = 0;
    int64_t count = 0;
    int kid = 1;
    int split = 32;
    int wgs_per_cu = 2;
    int grid = 0;
    int nitems = 0;
    int cus = 256;
    std::vector<int> h_n, h_nrhs;
    int* d_n = nullptr;
    int* d_nrhs = nullptr;
    int* d_order = nullptr;
    int4* d_items = nullptr;
    double tuned_ms = -1.0;
    char desc[256] = {0};
};

namespace tmvx {

// Build the work list for (kid, split): systems with n <= split in groups of
// (threads/64) per wave item, the rest one per group item; largest first.
static int plan_build(tmvx_batch_plan_s* p, int kid, int split, int wgs_per_cu, hipStream_t s)
{
    const BatchKernelInfo& K = kBatchKernels[kid];
    if (split > K.wave_nmax) split = K.wave_nmax;
    const int64_t cnt = p->count;
    std::vector<int> order((size_t)cnt);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return p->h_n[a] > p->h_n[b]; });
    std::vector<int4> items;
    const int per_wave_item = K.group_threads / 64;
    int64_t i = 0;
    for (; i < cnt && p->h_n[order[i]] > split; ++i) items.push_back(int4{0, (int)i, 1, 0});
    while (i < cnt) {
        const int c = (int)std::min<int64_t>(per_wave_item, cnt - i);
        items.push_back(int4{1, (int)i, c, 0});
        i += c;
    }
    p->kid = kid;
    p->split = split;
    p->wgs_per_cu = wgs_per_cu;
    p->nitems = (int)items.size();
    p->grid = std::max(1, std::min(p->nitems, p->cus * wgs_per_cu));
    if (p->d_items) { hipFree(p->d_items); p->d_items = nullptr; }
    if (!items.empty()) {
        TMVX_HIP(hipMalloc((void**)&p->d_items, sizeof(int4) * items.size()));
        TMVX_HIP(hipMemcpyAsync(p->d_items, items.data(), sizeof(int4) * items.size(),
                                hipMemcpyHostToDevice, s));
    }
    TMVX_HIP(hipMemcpyAsync(p->d_order, order.data(), sizeof(int) * (size_t)cnt,
                            hipMemcpyHostToDevice, s));
    TMVX_HIP(hipStreamSynchronize(s));
    snprintf(p->desc, sizeof(p->desc), "%s; split n<=%d per wave; %d items; grid %d (%d/CU)%s",
             K.name, split, p->nitems, p->grid, wgs_per_cu,
             p->tuned_ms >= 0 ? "; autotuned" : "");
    return TMVX_OK;
}

template <typename T>
static int batch_launch(const tmvx_batch_plan_s* p, T* const* A, i64* const* piv, T* const* B,
                        int* info, hipStream_t s)
{
    if (p->count == 0 || p->nitems == 0) return TMVX_OK;
    BatchArgs<T> args;
    args.order = p->d_order;
    args.items = p->d_items;
    args.nitems = p->nitems;
    args.n = p->d_n;
    args.nrhs = p->d_nrhs;
    args.A = A;
    args.piv = piv;
    args.B = B;
    args.info = info;
    const BatchKernelInfo& K = kBatchKernels[p->kid];
    const void* f = std::is_same<T, double>::value ? K.f64 : K.f32;
    void* kargs[] = {&args};
    TMVX_HIP(hipLaunchKernel(f, dim3((unsigned)p->grid), dim3((unsigned)K.group_threads), kargs,
                             0, s));
    return TMVX_OK;
}

// Synthetic batch of the plan's shape (uniform[-1,1] + n I, B uniform) for
// timing candidate schedules.
__global__ void batch_diag_shift_kernel(int count, const int* nv, double* const* A64,
                                        float* const* A32)
{
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const int n = nv[idx];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (A64) A64[idx][i + (i64)i * n] += (double)n;
        if (A32) A32[idx][i + (i64)i * n] += (float)n;
    }
}
__global__ void f64_to_f32_kernel(i64 cnt, const double* in, float* out)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) out[i] = (float)in[i];
}

static int plan_tune(tmvx_batch_plan_s* p, hipStream_t s)
{
    const int64_t cnt = p->count;
    std::vector<i64> offA((size_t)cnt), offB((size_t)cnt), offP((size_t)cnt);
    i64 ta = 0, tb = 0, tp = 0;
    for (int64_t i = 0; i < cnt; ++i) {
        offA[i] = ta; ta += (i64)p->h_n[i] * p->h_n[i];
        offB[i] = tb; tb += (i64)p->h_n[i] * p->h_nrhs[i];
        offP[i] = tp; tp += p->h_n[i];
    }
    const size_t esz = p->dtype == 0 ? 8 : 4;
    const i64 ta1 = ta > 0 ? ta : 1, tb1 = tb > 0 ? tb : 1, tp1 = tp > 0 ? tp : 1;
    char *dA0 = nullptr, *dA = nullptr, *dB0 = nullptr, *dB = nullptr;
    i64* dP = nullptr;
    void** dptr = nullptr;        // [A | B | piv | A0] pointer arrays
    int* dinfo = nullptr;
    double* stage = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TMVX_OK;
    if (hipMalloc((void**)&dA0, esz * ta1) != hipSuccess ||
        hipMalloc((void**)&dA, esz * ta1) != hipSuccess ||
        hipMalloc((void**)&dB0, esz * tb1) != hipSuccess ||
        hipMalloc((void**)&dB, esz * tb1) != hipSuccess ||
        hipMalloc((void**)&dP, sizeof(i64) * tp1) != hipSuccess ||
        hipMalloc((void**)&dptr, sizeof(void*) * 4 * cnt) != hipSuccess ||
        hipMalloc((void**)&dinfo, sizeof(int) * cnt) != hipSuccess ||
        hipMalloc((void**)&stage, sizeof(double) * std::max(ta1, tb1)) != hipSuccess)
        rc = TMVX_ERR_ALLOC;
    if (rc == TMVX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = TMVX_ERR_HIP;
    if (rc == TMVX_OK) {
        std::vector<void*> hp((size_t)(4 * cnt));
        for (int64_t i = 0; i < cnt; ++i) {
            hp[i] = dA + esz * offA[i];
            hp[cnt + i] = dB + esz * offB[i];
            hp[2 * cnt + i] = dP + offP[i];
            hp[3 * cnt + i] = dA0 + esz * offA[i];
        }
        if (hipMemcpy(dptr, hp.data(), sizeof(void*) * 4 * cnt, hipMemcpyHostToDevice) !=
            hipSuccess)
            rc = TMVX_ERR_HIP;
    }
    // Synthetic batch: uniform[-1,1] from a fixed seed, + n on the diagonal.
    auto fill = [&](uint64_t seed, i64 count, char* dst) -> int {
        if (count == 0) return TMVX_OK;
        double* tgt = p->dtype == 0 ? (double*)dst : stage;
        TMVX_TRY(tmvx_fill_uniform_f64(seed, 0, count, -1.0, 1.0, tgt, s));
        if (p->dtype == 1)
            hipLaunchKernelGGL(f64_to_f32_kernel, dim3((unsigned)((count + 255) / 256)),
                               dim3(256), 0, s, count, (const double*)stage, (float*)dst);
        return hip_status(hipGetLastError());
    };
    if (rc == TMVX_OK) rc = fill(0x7A15EED, ta, dA0);
    if (rc == TMVX_OK) rc = fill(0x7A15EEE, tb, dB0);
    if (rc == TMVX_OK && cnt > 0) {
        hipLaunchKernelGGL(batch_diag_shift_kernel, dim3((unsigned)cnt), dim3(128), 0, s,
                           (int)cnt, (const int*)p->d_n,
                           p->dtype == 0 ? (double* const*)(dptr + 3 * cnt) : nullptr,
                           p->dtype == 1 ? (float* const*)(dptr + 3 * cnt) : nullptr);
        rc = hip_status(hipGetLastError());
    }
    // Candidates: every schedule pair x split x workgroups per CU.
    struct Cand { int kid, split, wpc; };
    std::vector<Cand> cands;
    for (int k = 0; k < kNumBatchKernels; ++k)
        for (int sp : {0, 8, 16, 24, 32, 48, 64})
            if (sp <= kBatchKernels[k].wave_nmax)
                for (int w : {1, 2, 4}) cands.push_back({k, sp, w});
    double best = 1e30;
    Cand bestc = {p->kid, p->split, p->wgs_per_cu};
    for (size_t ci = 0; ci < cands.size() && rc == TMVX_OK; ++ci) {
        const Cand& c = cands[ci];
        bool ok = true;      // every system's RHS must fit the schedule that runs it
        for (int64_t i = 0; i < cnt && ok; ++i) {
            const bool wave = p->h_n[i] <= std::min(c.split, kBatchKernels[c.kid].wave_nmax);
            ok = p->h_nrhs[i] <= (wave ? kBatchKernels[c.kid].wave_rhsmax
                                       : kBatchKernels[c.kid].group_rhsmax);
        }
        if (!ok) continue;
        if ((rc = plan_build(p, c.kid, c.split, c.wpc, s)) != TMVX_OK) break;
        float msbest = 1e30f;
        for (int rep = 0; rep < 3 && rc == TMVX_OK; ++rep) {   // rep 0 warms up
            if (hipMemcpyAsync(dA, dA0, esz * ta, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipMemcpyAsync(dB, dB0, esz * tb, hipMemcpyDeviceToDevice, s) != hipSuccess) {
                rc = TMVX_ERR_HIP;
                break;
            }
            hipEventRecord(e0, s);
            if (p->dtype == 0)
                rc = batch_launch<double>(p, (double* const*)dptr, (i64* const*)(dptr + 2 * cnt),
                                          (double* const*)(dptr + cnt), dinfo, s);
            else
                rc = batch_launch<float>(p, (float* const*)dptr, (i64* const*)(dptr + 2 * cnt),
                                         (float* const*)(dptr + cnt), dinfo, s);
            hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) { rc = TMVX_ERR_HIP; break; }
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0) msbest = std::min(msbest, ms);
        }
        if (msbest < best) { best = msbest; bestc = c; }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(dA0); hipFree(dA); hipFree(dB0); hipFree(dB); hipFree(dP); hipFree(dptr);
    hipFree(dinfo); hipFree(stage);
    if (rc != TMVX_OK) return rc;
    p->tuned_ms = best;
    return plan_build(p, bestc.kid, bestc.split, bestc.wpc, s);
}

} // namespace tmvx

using namespace tmvx;

extern "C" {

int tmvx_batch_plan_create(int dtype, int64_t count, const int32_t* n_host,
                           const int32_t* nrhs_host, int tune, tmvx_batch_plan* plan,
                           void* stream)
{
    if (!plan) return TMVX_ERR_ARG;
    *plan = nullptr;
    if ((dtype != TMVX_BATCH_F64 && dtype != TMVX_BATCH_F32) || count < 0 || count > INT_MAX ||
        (count > 0 && !n_host))
        return TMVX_ERR_ARG;
    int maxn = 0, maxr = 0;
    for (int64_t i = 0; i < count; ++i) {
        const int ni = n_host[i], ri = nrhs_host ? nrhs_host[i] : 0;
        if (ni < 0 || ri < 0) return TMVX_ERR_ARG;
        maxn = std::max(maxn, ni);
        maxr = std::max(maxr, ri);
    }
    if (maxn > TMVX_BATCH_NMAX || maxr > TMVX_BATCH_RHSMAX) return TMVX_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    tmvx_batch_plan_s* p = new (std::nothrow) tmvx_batch_plan_s();
    if (!p) return TMVX_ERR_ALLOC;
    p->dtype = dtype;
    p->count = count;
    p->h_n.assign(n_host, n_host + count);
    p->h_nrhs.assign((size_t)count, 0);
    if (nrhs_host) p->h_nrhs.assign(nrhs_host, nrhs_host + count);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        p->cus = prop.multiProcessorCount;
    const size_t c1 = (size_t)std::max<int64_t>(count, 1);
    int rc = TMVX_OK;
    if (hipMalloc((void**)&p->d_n, sizeof(int) * c1) != hipSuccess ||
        hipMalloc((void**)&p->d_nrhs, sizeof(int) * c1) != hipSuccess ||
        hipMalloc((void**)&p->d_order, sizeof(int) * c1) != hipSuccess)
        rc = TMVX_ERR_ALLOC;
    if (rc == TMVX_OK && count > 0 &&
        (hipMemcpyAsync(p->d_n, p->h_n.data(), sizeof(int) * count, hipMemcpyHostToDevice, s) !=
             hipSuccess ||
         hipMemcpyAsync(p->d_nrhs, p->h_nrhs.data(), sizeof(int) * count, hipMemcpyHostToDevice,
                        s) != hipSuccess))
        rc = TMVX_ERR_HIP;
    // default schedule: waves for n <= 32, workgroups above, 2 per CU
    if (rc == TMVX_OK) rc = plan_build(p, 1, 32, 2, s);
    if (rc == TMVX_OK && tune && count > 0) rc = plan_tune(p, s);
    if (rc == TMVX_OK) {
        // the chosen schedule must hold every system's right-hand sides
        const BatchKernelInfo& K = kBatchKernels[p->kid];
        for (int64_t i = 0; i < count; ++i) {
            const bool wave = p->h_n[i] <= p->split;
            if (p->h_nrhs[i] > (wave ? K.wave_rhsmax : K.group_rhsmax)) {
                // fall back to the group schedule for every system
                rc = plan_build(p, p->kid, 0, p->wgs_per_cu, s);
                break;
            }
        }
    }
    if (rc != TMVX_OK) {
        tmvx_batch_plan_destroy(p);
        return rc;
    }
    *plan = p;
    return TMVX_OK;
}

int tmvx_batch_plan_destroy(tmvx_batch_plan p)
{
    if (!p) return TMVX_OK;
    hipFree(p->d_n);
    hipFree(p->d_nrhs);
    hipFree(p->d_order);
    hipFree(p->d_items);
    delete p;
    return TMVX_OK;
}

const char* tmvx_batch_plan_describe(tmvx_batch_plan p)
{
    return p ? p->desc : "";
}

int tmvx_batch_plan_set(tmvx_batch_plan p, int schedule, int split, int wgs_per_cu, void* stream)
{
    if (!p || schedule < 0 || schedule >= kNumBatchKernels || split < 0 || wgs_per_cu < 1 ||
        wgs_per_cu > 16)
        return TMVX_ERR_ARG;
    const BatchKernelInfo& K = kBatchKernels[schedule];
    const int sp = std::min(split, K.wave_nmax);
    for (int64_t i = 0; i < p->count; ++i)
        if (p->h_nrhs[i] > (p->h_n[i] <= sp ? K.wave_rhsmax : K.group_rhsmax))
            return TMVX_ERR_UNSUPPORTED;
    p->tuned_ms = -1.0;
    return plan_build(p, schedule, sp, wgs_per_cu, (hipStream_t)stream);
}

int tmvx_batch_schedule_count(void) { return kNumBatchKernels; }

int tmvx_gesv_batched_f64(tmvx_batch_plan p, double* const* A, int64_t* const* piv,
                          double* const* B, int32_t* info, void* stream)
{
    if (!p || p->dtype != TMVX_BATCH_F64) return TMVX_ERR_ARG;
    if (p->count > 0 && (!A || !piv || !info)) return TMVX_ERR_ARG;
    return batch_launch<double>(p, A, (i64* const*)piv, B, info, (hipStream_t)stream);
}

int tmvx_gesv_batched_f32(tmvx_batch_plan p, float* const* A, int64_t* const* piv,
                          float* const* B, int32_t* info, void* stream)
{
    if (!p || p->dtype != TMVX_BATCH_F32) return TMVX_ERR_ARG;
    if (p->count > 0 && (!A || !piv || !info)) return TMVX_ERR_ARG;
    return batch_launch<float>(p, A, (i64* const*)piv, B, info, (hipStream_t)stream);
}

static int getrs_batched_launch(int trans, int64_t count, const int32_t* n, const int32_t* nrhs,
                                const void* LU, const int64_t* const* piv, void* B, int f64,
                                hipStream_t s)
{
    if (count < 0 || count > INT_MAX) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    if (!n || !nrhs || !LU || !piv || !B) return TMVX_ERR_ARG;
    const unsigned grid = 2048;      // 8192 waves; wave w takes systems w, w + 8192, ...
    if (f64)
        hipLaunchKernelGGL(batched_getrs_kernel<double>, dim3(grid), dim3(256), 0, s, trans,
                           (int)count, n, nrhs, (const double* const*)LU, (const i64* const*)piv,
                           (double* const*)B);
    else
        hipLaunchKernelGGL(batched_getrs_kernel<float>, dim3(grid), dim3(256), 0, s, trans,
                           (int)count, n, nrhs, (const float* const*)LU, (const i64* const*)piv,
                           (float* const*)B);
    return hip_status(hipGetLastError());
}

int tmvx_getrs_batched_f64(int trans, int64_t count, const int32_t* n, const int32_t* nrhs,
                           const double* const* LU, const int64_t* const* piv, double* const* B,
                           void* stream)
{
    return getrs_batched_launch(trans, count, n, nrhs, LU, piv, (void*)B, 1, (hipStream_t)stream);
}

int tmvx_getrs_batched_f32(int trans, int64_t count, const int32_t* n, const int32_t* nrhs,
                           const float* const* LU, const int64_t* const* piv, float* const* B,
                           void* stream)
{
    return getrs_batched_launch(trans, count, n, nrhs, LU, piv, (void*)B, 0, (hipStream_t)stream);
}

int tmvx_gemm_batched_f64(int64_t count, const int32_t* m, const int32_t* n, const int32_t* k,
                          double alpha, const double* const* A, const double* const* B,
                          double beta, double* const* C, void* stream)
{
    if (count < 0 || count > INT_MAX) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    if (!m || !n || !k || !A || !B || !C) return TMVX_ERR_ARG;
    const unsigned grid = (unsigned)std::min<int64_t>(count, 2048);
    hipLaunchKernelGGL(batched_gemm_kernel<double>, dim3(grid), dim3(256), 0,
                       (hipStream_t)stream, (int)count, m, n, k, alpha, A, B, beta, C);
    return hip_status(hipGetLastError());
}

int tmvx_gemm_batched_f32(int64_t count, const int32_t* m, const int32_t* n, const int32_t* k,
                          float alpha, const float* const* A, const float* const* B, float beta,
                          float* const* C, void* stream)
{
    if (count < 0 || count > INT_MAX) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    if (!m || !n || !k || !A || !B || !C) return TMVX_ERR_ARG;
    const unsigned grid = (unsigned)std::min<int64_t>(count, 2048);
    hipLaunchKernelGGL(batched_gemm_kernel<float>, dim3(grid), dim3(256), 0,
                       (hipStream_t)stream, (int)count, m, n, k, alpha, A, B, beta, C);
    return hip_status(hipGetLastError());
}

} // extern "C"
