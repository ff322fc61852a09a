// Tournament pivoting (TSLU, the panel factorisation of communication-
// avoiding LU) for FP32 factorisations that do not need TMV's pivot
// sequence (the mixed-precision solver: iterative refinement in FP64
// corrects the solution).  It trades the per-column grid-wide exchange of
// partial pivoting for a few kernel boundaries per panel.
//
// Panel [j0, j0 + jb) of the m = n - j0 rows below j0:
//   1. select: workgroups of 256 threads (one wave per SIMD) each take up
//      to 256 candidate rows (one per thread, column c in register c) and
//      run partial pivoting among them.  Per column ONE barrier: every wave
//      finds its best candidate as the maximum of a 32-bit key (|value| and
//      thread index packed) by DPP, and that lane writes its row into the
//      wave's LDS slot (double-buffered by column parity); after the
//      barrier every thread takes the largest of the 4 wave keys (the key
//      names the winning thread), reads the pivot row and updates its own
//      row's remaining columns.  Level 1 takes the panel's rows in slabs of
//      256; each further level takes 4 previous outputs (256 rows) per
//      workgroup, until one workgroup remains (m = 16384: 64, 16, 4, 1
//      workgroups).  Candidates are always the original rows of A.
//   2. the last level (one workgroup) has factored the chosen jb x jb block
//      with partial pivoting: it writes L11\U11, the interchange list
//      piv[j0 + j] (LAPACK / TMV semantics, by simulating the swaps on a
//      table of at most 2 jb displaced rows held in lanes), where each
//      displaced row ends up, and a copy of the panel's original top jb rows.
//   3. factor: every workgroup loads U11; workgroup 0 writes L11\U11 over
//      the top block; every thread takes one row below it -- the displaced
//      top row (from the copy) if the interchanges moved one there, else its
//      own -- and solves x U11 = a (L21 = A21 U11^-1).  The panel's own
//      interchanges need no separate pass.  A pivot below FLT_MIN in U11
//      (singular in the working precision) is reported as the first zero
//      pivot and its column of L is zeroed, as TMV does for partial pivoting.
//
// Measured (FP32 LU, n = 16384, with the look-ahead driver): 121 ms against
// 125 ms with the partial-pivoting panel.  One selection level costs
// ~3k cycles per column (key maximum, LDS row exchange, barrier, update),
// about what the partial-pivoting panel spends per column including its
// grid exchange, and the tree needs 4 levels; variants with one wave per
// workgroup (no barrier, 2-3 rows per lane) or a runtime column loop were
// slower.  The mixed-precision solver therefore keeps partial pivoting.
#include "tmvx_lu.h"
#include "tmvx_pivot.h"

#include <climits>

namespace tmvx {

constexpr int TS_NB = 64;      // panel width
constexpr int TS_NTH = 256;    // rows per workgroup (one per thread, 4 waves)

// -DTMVX_TSLU_TRACE: clock64 per phase of each column, workgroup 0,
// thread 0, first-level selection of the first panel.
#ifdef TMVX_TSLU_TRACE
__device__ unsigned long long* g_tslu_trace;
#define TSLU_TRACE(k, ph)                                                      \
    do {                                                                       \
        if (!FINAL && g_tslu_trace && w == 0 && t == 0 && j0 == 0 && !in_idx)  \
            g_tslu_trace[(k) * 8 + (ph)] = clock64();                          \
    } while (0)
#else
#define TSLU_TRACE(k, ph) do {} while (0)
#endif

// Workspace written by the last level (after the level outputs).
struct TsluFinal {
    float* lu;      // NB x NB row-major: packed L11\U11, row k = k-th pivot row
    float* top;     // NB x NB row-major: the panel's original top rows
    int* map_pos;   // positions (panel-relative) whose row changed ...
    int* map_src;   // ... and the original panel-relative row now there
    int* map_cnt;
    i64* piv;
};

// Interchange list from the winners (panel-relative rows, pivot order):
// position k initially holds row k; swapping positions j and p is recorded
// in a table of (position, row) pairs for the positions that changed
// (lane l holds entries l and l + 64).  One wave.
__device__ __forceinline__ void tslu_make_piv(const int* s_win, i64 j0, int jb, i64 m,
                                              const TsluFinal& fin)
{
    const int lane = threadIdx.x & 63;
    int tpos[2] = {-1, -1}, trow[2] = {-1, -1};
    int cnt = 0;
    auto find_row = [&](int r) __attribute__((always_inline)) -> int {   // position of row r
        int p = r;
        const unsigned long long h0 = __ballot(trow[0] == r), h1 = __ballot(trow[1] == r);
        if (h0) p = __builtin_amdgcn_readlane(tpos[0], __builtin_ctzll(h0));
        else if (h1) p = __builtin_amdgcn_readlane(tpos[1], __builtin_ctzll(h1));
        return p;
    };
    auto row_at = [&](int p) __attribute__((always_inline)) -> int {     // row at position p
        int r = p;
        const unsigned long long h0 = __ballot(tpos[0] == p), h1 = __ballot(tpos[1] == p);
        if (h0) r = __builtin_amdgcn_readlane(trow[0], __builtin_ctzll(h0));
        else if (h1) r = __builtin_amdgcn_readlane(trow[1], __builtin_ctzll(h1));
        return r;
    };
    auto set = [&](int p, int r) __attribute__((always_inline)) {         // position p := row r
        const unsigned long long h0 = __ballot(tpos[0] == p), h1 = __ballot(tpos[1] == p);
        if (h0) {
            if (lane == __builtin_ctzll(h0)) trow[0] = r;
        } else if (h1) {
            if (lane == __builtin_ctzll(h1)) trow[1] = r;
        } else {
            const int slot = cnt >> 6, l = cnt & 63;
            if (lane == l) { if (slot == 0) { tpos[0] = p; trow[0] = r; } else { tpos[1] = p; trow[1] = r; } }
            ++cnt;
        }
    };
    for (int j = 0; j < jb; ++j) {
        int target = s_win[j];
        if (target < 0 || target >= m) target = row_at(j);       // defensive: keep row j
        const int p = find_row(target);
        if (lane == 0) fin.piv[j0 + j] = j0 + p;
        if (p != j) {
            const int rj = row_at(j);
            set(j, target);
            set(p, rj);
        }
    }
    if (lane < cnt) { fin.map_pos[lane] = tpos[0]; fin.map_src[lane] = trow[0]; }
    if (lane + 64 < cnt) { fin.map_pos[lane + 64] = tpos[1]; fin.map_src[lane + 64] = trow[1]; }
    if (lane == 0) *fin.map_cnt = cnt;
}

// Wave maximum of an unsigned key by DPP (quad swaps, half-row and row
// mirrors, row broadcasts); every lane gets lane 63's result.
__device__ __forceinline__ unsigned wave_max_u32(unsigned k)
{
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0xB1, 0xf, 0xf, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x4E, 0xf, 0xf, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x141, 0xf, 0xf, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x140, 0xf, 0xf, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x142, 0xa, 0xf, false));
    k = max(k, (unsigned)__builtin_amdgcn_update_dpp((int)k, (int)k, 0x143, 0xc, 0xf, false));
    return (unsigned)__builtin_amdgcn_readlane((int)k, 63);
}

// Partial pivoting among up to TS_NTH listed rows (one per thread, column
// c of the panel in register c; the column loop is unrolled so every
// register index is static and the update covers only the triangle
// c > j).  Candidates are ranked by a 32-bit key: |value|'s bits with the
// low 9 mantissa bits replaced by 511 - thread, so one unsigned maximum
// finds the largest |value| to 2^-14 relative (ties and near-ties to the
// lowest thread) and names the winning thread; the pivot is any candidate
// within that margin of the largest, which partial pivoting's stability
// does not notice.  An updated row keeps its multiplier in the register of
// the column just eliminated, so a chosen row's registers are its packed
// L\U row.
// Non-final levels write the jb chosen panel-relative row indices (pivot
// order, -1 when a short slab runs out) to out[blockIdx.x * jb ...]; the
// final level (one workgroup) writes L11\U11, piv, the displacement map and
// the copy of the top rows.
template <bool FINAL>
__global__ void __launch_bounds__(TS_NTH)
tslu_select_kernel(const float* __restrict__ A, i64 lda, i64 j0, int jb, int total, i64 m,
                   const int* __restrict__ in_idx, int* __restrict__ out, TsluFinal fin)
{
    typedef float T;
    constexpr int NB = TS_NB, NTH = TS_NTH, NW = NTH / 64;
    typedef typename Vec16<T>::type VT;
    constexpr int VN = Vec16<T>::n, NV = NB / VN;
    static_assert(NW == 4 && NTH <= 256, "key layout: 4 waves, thread index in 8 bits");
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, w = blockIdx.x;
    __shared__ VT s_rows[2][NW][NV];                       // wave candidates, by column parity
    __shared__ __attribute__((aligned(16))) unsigned s_key[2][NW];
    __shared__ int s_idx[NTH];
    __shared__ int s_win[NB];
    __shared__ VT s_lu[FINAL ? NB : 1][NV];                 // chosen rows (packed L\U rows)

    const int e = w * NTH + t;
    const int idx = e < total ? (in_idx ? in_idx[e] : e) : -1;   // -1: no candidate
    s_idx[t] = idx;
    T row[NB];
    {
        const unsigned off = idx >= 0 ? (unsigned)idx : 0u;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const T* colp = A + j0 + (j0 + (i64)(c < jb ? c : jb - 1)) * lda;
            const T v = colp[off];
            row[c] = (idx >= 0 && c < jb) ? v : T(0);
        }
    }
    bool active = idx >= 0;
    const unsigned tkey = 511u - (unsigned)t;              // >= 256: a candidate's key is never 0
    __syncthreads();

    static_for<0, NB>([&](auto jj) __attribute__((always_inline)) {
        constexpr int j = decltype(jj)::value;
        constexpr int g0 = j / VN;                          // first vector holding column j
        if (j < jb) {
            constexpr int par = j & 1;
            TSLU_TRACE(j, 0);
            const unsigned key =
                active ? ((__float_as_uint(fabsf(row[j])) & 0xFFFFFE00u) | tkey) : 0u;
            const unsigned wkey = wave_max_u32(key);
            TSLU_TRACE(j, 1);
            if (active && key == wkey) {
                static_for<g0, NV>([&](auto gg) __attribute__((always_inline)) {
                    constexpr int g = decltype(gg)::value;
                    VT x;
                    static_for<0, VN>([&](auto u) __attribute__((always_inline)) {
                        x[decltype(u)::value] = row[g * VN + decltype(u)::value];
                    });
                    s_rows[par][wave][g] = x;
                });
            }
            if (lane == 0) s_key[par][wave] = wkey;
            TSLU_TRACE(j, 2);
            __syncthreads();
            TSLU_TRACE(j, 3);
            const uint4 k4 = *reinterpret_cast<const uint4*>(&s_key[par][0]);
            const unsigned best = max(max(k4.x, k4.y), max(k4.z, k4.w));
            const int tw = 511 - (int)(best & 511u);       // winning thread (if best != 0)
            const bool none = best == 0u;
            const bool zero = (best & 0xFFFFFE00u) < 0x00800000u;   // |pivot| < FLT_MIN
            if (t == 0) s_win[j] = none ? -1 : s_idx[tw];
            TSLU_TRACE(j, 4);
            if (!none && t == tw) {
                if constexpr (FINAL) {
                    static_for<0, NV>([&](auto gg) __attribute__((always_inline)) {
                        constexpr int g = decltype(gg)::value;
                        VT x;
                        static_for<0, VN>([&](auto u) __attribute__((always_inline)) {
                            x[decltype(u)::value] = row[g * VN + decltype(u)::value];
                        });
                        s_lu[j][g] = x;
                    });
                }
                active = false;
            }
            if (active) {
                const int ww = tw >> 6;
                const VT pj = s_rows[par][ww][g0];
                const T l = zero ? T(0) : row[j] * __builtin_amdgcn_rcpf(pj[j % VN]);
                row[j] = l;
                static_for<g0, NV>([&](auto gg) __attribute__((always_inline)) {
                    constexpr int g = decltype(gg)::value;
                    const VT x = g == g0 ? pj : s_rows[par][ww][g];
                    static_for<0, VN>([&](auto u) __attribute__((always_inline)) {
                        constexpr int c = g * VN + decltype(u)::value;
                        if constexpr (c > j) row[c] -= l * x[decltype(u)::value];
                    });
                });
            }
            TSLU_TRACE(j, 5);
        }
    });
    __syncthreads();
    if constexpr (!FINAL) {
        if (t < jb) out[(i64)w * jb + t] = s_win[t];
    } else {
        if (wave == 0) tslu_make_piv(s_win, j0, jb, m, fin);
        const T* slu = reinterpret_cast<const T*>(&s_lu[0][0]);
        for (int q = t; q < NB * NB; q += NTH) {
            const int k = q / NB, col = q % NB;
            fin.lu[q] = (k < jb && col < jb) ? slu[q] : T(0);
        }
        // the panel's original top rows (read before anything rewrites them)
        for (int q = t; q < NB * NB; q += NTH) {
            const int r = q % NB, c = q / NB;
            fin.top[r * NB + c] = (r < jb && c < jb) ? A[(j0 + r) + (j0 + c) * lda] : T(0);
        }
    }
}

// The panel's L and U once the pivots are known: L11\U11 from the last
// level; rows below: the interchanges' result (own row or displaced top
// row), then x U11 = a.
__global__ void __launch_bounds__(TS_NTH)
tslu_factor_kernel(float* __restrict__ A, i64 lda, i64 j0, int jb, i64 m, TsluFinal fin,
                   i64* __restrict__ first_zero)
{
    typedef float T;
    constexpr int NB = TS_NB, NTH = TS_NTH;
    typedef typename Vec16<T>::type VT;
    constexpr int VN = Vec16<T>::n, NV = NB / VN;
    const int t = threadIdx.x, w = blockIdx.x;
    __shared__ VT s_u[NB][NV];               // L11\U11 row-major
    __shared__ VT s_ur[NB][NV];              // row k of U11 shifted left by k (s_ur[k][c] = U[k][k+c])
    __shared__ T s_rd[NB];                   // reciprocals of U11's diagonal (0 if tiny)
    __shared__ int s_src[NTH];               // displaced top row landing in this slab, or -1
    {
        const VT* g = reinterpret_cast<const VT*>(fin.lu);
        VT* su = &s_u[0][0];
        for (int q = t; q < NB * NV; q += NTH) su[q] = g[q];
    }
    s_src[t] = -1;
    __syncthreads();
    {
        const T* su = reinterpret_cast<const T*>(&s_u[0][0]);
        T* sr = reinterpret_cast<T*>(&s_ur[0][0]);
        for (int q = t; q < NB * NB; q += NTH) {
            const int k = q / NB, c = q % NB;
            sr[q] = (k + c < NB) ? su[k * NB + k + c] : T(0);
        }
    }
    const i64 base = (i64)jb + (i64)w * NTH;   // panel-relative row of thread 0
    {
        const int cnt = *fin.map_cnt;
        for (int q = t; q < cnt; q += NTH) {
            const i64 pos = fin.map_pos[q];
            const int src = fin.map_src[q];
            if (pos >= base && pos < base + NTH && src >= 0 && src < jb) s_src[pos - base] = src;
        }
    }
    if (t < NB) {
        const T d = reinterpret_cast<const T*>(&s_u[t][0])[t];
        const bool tiny_d = !(fabsf(d) >= tiny<T>());
        s_rd[t] = tiny_d ? T(0) : T(1) / d;
        if (tiny_d && t < jb && w == 0)
            atomicMin(reinterpret_cast<unsigned long long*>(first_zero),
                      (unsigned long long)(j0 + t + 1));
    }
    __syncthreads();
    if (w == 0) {
        const T* su = reinterpret_cast<const T*>(&s_u[0][0]);
        for (int q = t; q < jb * jb; q += NTH) {
            const int r = q % jb, c = q / jb;
            A[(j0 + r) + (j0 + c) * lda] = su[r * NB + c];
        }
    }
    const i64 r = base + t;
    if (r < m) {
        T* gp = A + (j0 + r) + j0 * lda;
        const int src = s_src[t];
        // own row (columns clamped into the panel, padding selected to 0
        // after the load: uniform offsets, no branches); a displaced top row
        // (at most jb threads) replaces it from the copy
        T x[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const T v = gp[(i64)(c < jb ? c : jb - 1) * lda];
            x[c] = c < jb ? v : T(0);
        }
        if (src >= 0) {
            const VT* tp = reinterpret_cast<const VT*>(fin.top + (i64)src * NB);
            static_for<0, NV>([&](auto q) __attribute__((always_inline)) {
                const VT y = tp[decltype(q)::value];
                static_for<0, VN>([&](auto u) __attribute__((always_inline)) {
                    x[decltype(q)::value * VN + decltype(u)::value] = y[decltype(u)::value];
                });
            });
        }
        // x U11 = a column by column, x rotated so the current column is
        // x[0]: the solved value goes to the top as the rotation frees it
        // (a runtime loop: a fully unrolled triangle keeps every U11 value
        // live at once); columns >= jb have rd = 0 and zero U rows.
        for (int k = 0; k < NB; ++k) {
            const T xk = x[0] * s_rd[k];
            VT u[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) u[q] = s_ur[k][q];
            x[0] = xk;
#pragma unroll
            for (int c = 1; c < NB; ++c) x[c] -= xk * u[c / VN][c % VN];
            const T first = x[0];
#pragma unroll
            for (int c = 0; c < NB - 1; ++c) x[c] = x[c + 1];
            x[NB - 1] = first;
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < jb) gp[(i64)c * lda] = x[c];
    }
}

i64 tslu_ws_ints(i64 m)
{
    const i64 g = (m + TS_NTH - 1) / TS_NTH;
    // level outputs | (alignment) lu | top | map_pos | map_src | map_cnt
    return 2 * g * TS_NB + 64 + 2 * TS_NB * TS_NB + 2 * (2 * TS_NB) + 64;
}

// Tournament-pivoted factorisation of panel [j0, j0+jb): selection levels,
// the last of which also forms L11\U11 and piv, then the factor kernel.
// ws: at least tslu_ws_ints(n - j0) ints of device workspace.
template <typename T>
int tslu_panel(T* A, i64 lda, i64 n, i64 j0, int jb, i64* piv, i64* first_zero, int* ws,
               hipStream_t s)
{
    static_assert(sizeof(T) == 4, "FP32 only");
    const i64 m = n - j0;
    if (jb <= 0 || jb > TS_NB || m < jb || !ws) return TMVX_ERR_ARG;
    const i64 g1 = (m + TS_NTH - 1) / TS_NTH;
    int* lvl_a = ws;
    int* lvl_b = ws + g1 * TS_NB;
    int* tail = ws + 2 * g1 * TS_NB;
    tail += (16 - ((reinterpret_cast<uintptr_t>(tail) & 63) >> 2)) & 15;   // 64-byte align
    TsluFinal fin;
    fin.lu = reinterpret_cast<float*>(tail);
    fin.top = fin.lu + TS_NB * TS_NB;
    fin.map_pos = reinterpret_cast<int*>(fin.top + TS_NB * TS_NB);
    fin.map_src = fin.map_pos + 2 * TS_NB;
    fin.map_cnt = fin.map_src + 2 * TS_NB;
    fin.piv = piv;
    int total = (int)m;
    const int* in = nullptr;
    int* out = lvl_a;
    while (true) {
        const int g = (total + TS_NTH - 1) / TS_NTH;
        if (g == 1) {
            hipLaunchKernelGGL(tslu_select_kernel<true>, dim3(1), dim3(TS_NTH), 0, s,
                               (const float*)A, lda, j0, jb, total, m, in, out, fin);
            TMVX_CHECK_LAUNCH();
            break;
        }
        hipLaunchKernelGGL(tslu_select_kernel<false>, dim3((unsigned)g), dim3(TS_NTH), 0, s,
                           (const float*)A, lda, j0, jb, total, m, in, out, fin);
        TMVX_CHECK_LAUNCH();
        total = g * jb;
        in = out;
        out = (out == lvl_a) ? lvl_b : lvl_a;
    }
    const i64 below = m - jb;
    const unsigned gf = (unsigned)(below > 0 ? (below + TS_NTH - 1) / TS_NTH : 1);
    hipLaunchKernelGGL(tslu_factor_kernel, dim3(gf), dim3(TS_NTH), 0, s, (float*)A, lda, j0, jb,
                       m, fin, first_zero);
    return hip_status(hipGetLastError());
}

template int tslu_panel<float>(float*, i64, i64, i64, int, i64*, i64*, int*, hipStream_t);

} // namespace tmvx

#ifdef TMVX_TSLU_TRACE
extern "C" int tmvx_tslu_trace_set(unsigned long long* buf)
{
    return hipMemcpyToSymbol(HIP_SYMBOL(tmvx::g_tslu_trace), &buf, sizeof(buf)) == hipSuccess
               ? 0 : -3;
}
#endif
