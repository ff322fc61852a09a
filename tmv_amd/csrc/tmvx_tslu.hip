// Tournament pivoting (TSLU, the panel factorisation of communication-
// avoiding LU) for FP32 factorisations that do not need TMV's pivot
// sequence (the mixed-precision solver could use it: iterative refinement
// in FP64 corrects the solution).  It trades the per-column grid-wide
// exchange of partial pivoting for a few kernel boundaries per panel.
// Measured at n = 16384 it is slower than the partial-pivoting panel (160
// vs 127 ms per factorisation): one selection level costs ~120 us, i.e. the
// local per-column work alone is ~1.9 us, about what the partial-pivoting
// panel spends per column including its exchange.  Exposed through
// tmvx_getrf_f32_tournament only.
//
// Panel [j0, j0 + jb) of the m = n - j0 rows below j0:
//   1. select: workgroups of 512 threads each take up to 512 candidate
//      rows (one per thread, the panel's jb values in registers, rotated so
//      the current column is register 0) and run partial pivoting among
//      them -- argmax by DPP within waves and across the 8 waves through
//      LDS, pivot row broadcast through LDS, rank-1 update of the rows not
//      yet chosen -- and output the jb rows chosen, in order.  Only barriers,
//      no inter-workgroup communication.  Level 1 takes the panel's rows in
//      slabs of 512; each further level takes 8 previous outputs (512 rows)
//      per workgroup, until one set of jb rows remains (m = 16384: 32, 4, 1
//      workgroups).  Candidates are always the original rows of A.
//   2. piv: one wave turns the winners into an interchange list piv[j0 + j]
//      (LAPACK / TMV semantics), simulating the swaps on a table of at most
//      2 jb displaced rows held in lanes.
//   3. the panel's interchanges are applied to its own columns (the laswp
//      kernel of tmvx_lu.hip), then factor: every workgroup factors the
//      jb x jb top block without pivoting in LDS (redundantly), and solves
//      x U11 = a for each of its rows below (L21 = A21 U11^-1); workgroup 0
//      also writes L11\U11.  A pivot below FLT_MIN in U11 (singular in the
//      working precision) is reported as the first zero pivot and its
//      column of L is zeroed, as TMV does for partial pivoting.
#include "tmvx_lu.h"
#include "tmvx_pivot.h"

#include <climits>

namespace tmvx {

constexpr int TS_NB = 64;      // panel width
constexpr int TS_NTH = 512;    // candidate rows per workgroup

// Local partial pivoting among up to TS_NTH listed rows; writes the jb
// chosen panel-relative row indices (in pivot order) to out[blockIdx.x].
template <typename T>
__global__ void __launch_bounds__(TS_NTH)
tslu_select_kernel(const T* __restrict__ A, i64 lda, i64 j0, int jb, int total,
                   const int* __restrict__ in_idx, int* __restrict__ out)
{
    constexpr int NB = TS_NB, NTH = TS_NTH, NW = NTH / 64;
    typedef typename Vec16<T>::type VT;
    constexpr int VN = Vec16<T>::n;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, w = blockIdx.x;
    __shared__ __attribute__((aligned(16))) T s_prow[NB];
    __shared__ T s_wv[NW];
    __shared__ int s_wi[NW];
    __shared__ int s_win;
    __shared__ T s_bv;

    const int e = w * NTH + t;
    int idx = e < total ? (in_idx ? in_idx[e] : e) : -1;
    const bool valid = idx >= 0;                     // -1: no candidate (short slab)
    if (!valid) idx = INT_MAX;
    T row[NB];
    const T* gp = A + (j0 + (valid ? idx : 0)) + j0 * lda;
#pragma unroll
    for (int c = 0; c < NB; ++c) row[c] = (valid && c < jb) ? gp[(i64)c * lda] : T(0);
    bool active = valid;

    for (int j = 0; j < jb; ++j) {
        T v = active ? fabs(row[0]) : T(-1);
        int vi = active ? idx : INT_MAX;
        wave_argmax(v, vi);
        if (lane == 0) { s_wv[wave] = v; s_wi[wave] = vi; }
        __syncthreads();
        if (wave == 0) {
            T bv = lane < NW ? s_wv[lane] : T(-1);
            int bi = lane < NW ? s_wi[lane] : INT_MAX;
            wave_argmax(bv, bi);
            if (lane == 0) { s_bv = bv; s_win = bi; }
        }
        __syncthreads();
        const int win = s_win;
        const bool zero = !(s_bv >= tiny<T>());
        if (active && idx == win) {
            static_for<0, NB / VN>([&](auto cc) {
                constexpr int c = decltype(cc)::value * VN;
                VT x;
                static_for<0, VN>([&](auto uc) { x[decltype(uc)::value] = row[c + decltype(uc)::value]; });
                *reinterpret_cast<VT*>(&s_prow[c]) = x;
            });
            active = false;
        }
        if (t == 0) out[(i64)w * jb + j] = win == INT_MAX ? -1 : win;
        __syncthreads();
        if (active && !zero) {
            const T l = row[0] * (T(1) / s_prow[0]);
#pragma unroll
            for (int c = 1; c < NB; ++c)
                if (c < jb - j) row[c] -= l * s_prow[c];
        }
        // rotate: the next column into register 0
        const T first = row[0];
#pragma unroll
        for (int c = 0; c < NB - 1; ++c) row[c] = row[c + 1];
        row[NB - 1] = first;
        __syncthreads();            // s_prow and s_wv are rewritten next column
    }
}

// Winners (panel-relative rows, pivot order) -> interchange list.  One wave.
// Position k initially holds row k; swapping positions j and p is recorded
// in a table of (position, row) pairs for the positions that changed.
__global__ void __launch_bounds__(64)
tslu_piv_kernel(i64 j0, int jb, i64 m, const int* __restrict__ win, i64* __restrict__ piv)
{
    const int lane = threadIdx.x;
    int tpos[2] = {-1, -1}, trow[2] = {-1, -1};     // lane l holds entries l, l + 64
    int cnt = 0;
    auto find_row = [&](int r) -> int {            // current position of row r
        int p = r;
        const unsigned long long h0 = __ballot(trow[0] == r), h1 = __ballot(trow[1] == r);
        if (h0) p = __builtin_amdgcn_readlane(tpos[0], __builtin_ctzll(h0));
        else if (h1) p = __builtin_amdgcn_readlane(tpos[1], __builtin_ctzll(h1));
        return p;
    };
    auto row_at = [&](int p) -> int {              // row now at position p
        int r = p;
        const unsigned long long h0 = __ballot(tpos[0] == p), h1 = __ballot(tpos[1] == p);
        if (h0) r = __builtin_amdgcn_readlane(trow[0], __builtin_ctzll(h0));
        else if (h1) r = __builtin_amdgcn_readlane(trow[1], __builtin_ctzll(h1));
        return r;
    };
    auto set = [&](int p, int r) {                 // position p now holds row r
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
        int target = win[j];
        if (target < 0 || target >= m) target = row_at(j);      // defensive: keep row j
        const int p = find_row(target);
        if (lane == 0) piv[j0 + j] = j0 + p;
        if (p != j) {
            const int rj = row_at(j);
            set(j, target);
            set(p, rj);
        }
    }
}

// Factor the panel with its pivots applied: U11 and L11 from the top block
// (no pivoting, in LDS), L21 = A21 U11^-1 one row per thread.
template <typename T>
__global__ void __launch_bounds__(TS_NTH)
tslu_factor_kernel(T* __restrict__ A, i64 lda, i64 j0, int jb, i64 m, i64* __restrict__ first_zero)
{
    constexpr int NB = TS_NB, NTH = TS_NTH;
    const int t = threadIdx.x, w = blockIdx.x;
    __shared__ T s_u[NB][NB + 1];            // top block, factored in place
    __shared__ T s_rd[NB];                   // reciprocals of the diagonal (0 if tiny)
    for (int e = t; e < NB * NB; e += NTH) {
        const int r = e % NB, c = e / NB;
        s_u[r][c] = (r < jb && c < jb) ? A[(j0 + r) + (j0 + c) * lda] : T(0);
    }
    __syncthreads();
    for (int k = 0; k < jb; ++k) {
        const T d = s_u[k][k];
        const bool tiny_d = !(fabs(d) >= tiny<T>());
        const T rd = tiny_d ? T(0) : T(1) / d;
        // column k below the diagonal: multipliers; then the rank-1 update
        for (int e = t; e < NB * NB; e += NTH) {
            const int r = e % NB, c = e / NB;
            if (r > k && r < jb && c > k && c < jb) s_u[r][c] -= (s_u[r][k] * rd) * s_u[k][c];
        }
        __syncthreads();
        if (t > k && t < jb) s_u[t][k] = s_u[t][k] * rd;
        if (t == 0) {
            s_rd[k] = rd;
            if (tiny_d && w == 0)
                atomicMin(reinterpret_cast<unsigned long long*>(first_zero),
                          (unsigned long long)(j0 + k + 1));
        }
        __syncthreads();
    }
    if (w == 0) {
        for (int e = t; e < jb * jb; e += NTH) {
            const int r = e % jb, c = e / jb;
            A[(j0 + r) + (j0 + c) * lda] = s_u[r][c];
        }
    }
    // rows below the top block: x U11 = a
    const i64 r = (i64)jb + (i64)w * NTH + t;        // panel-relative row
    if (r < m) {
        T* gp = A + (j0 + r) + j0 * lda;
        T x[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) x[c] = c < jb ? gp[(i64)c * lda] : T(0);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k < jb) {
                const T xk = x[k] * s_rd[k];
                x[k] = xk;
#pragma unroll
                for (int c = k + 1; c < NB; ++c) x[c] -= xk * s_u[k][c];
            }
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < jb) gp[(i64)c * lda] = x[c];
    }
}

// Tournament-pivoted factorisation of panel [j0, j0+jb): selection tree,
// interchange list, interchanges on the panel's columns, factor.
// ws: at least tslu_ws_ints(m) ints of device workspace.
template <typename T>
int tslu_panel(T* A, i64 lda, i64 n, i64 j0, int jb, i64* piv, i64* first_zero, int* ws,
               hipStream_t s)
{
    const i64 m = n - j0;
    int total = (int)m;
    const int* in = nullptr;
    int* out = ws;
    int* other = ws + ((m + TS_NTH - 1) / TS_NTH) * TS_NB;
    while (true) {
        const int g = (total + TS_NTH - 1) / TS_NTH;
        hipLaunchKernelGGL(tslu_select_kernel<T>, dim3((unsigned)g), dim3(TS_NTH), 0, s, A, lda, j0,
                           jb, total, in, out);
        TMVX_CHECK_LAUNCH();
        if (g == 1) break;
        total = g * jb;
        in = out;
        int* nxt = (out == ws) ? other : ws;
        out = nxt;
    }
    hipLaunchKernelGGL(tslu_piv_kernel, dim3(1), dim3(64), 0, s, j0, jb, m, (const int*)out, piv);
    TMVX_CHECK_LAUNCH();
    return TMVX_OK;
}

i64 tslu_ws_ints(i64 m)
{
    const i64 g = (m + TS_NTH - 1) / TS_NTH;
    return 2 * g * TS_NB + 64;
}

template <typename T>
int tslu_factor(T* A, i64 lda, i64 n, i64 j0, int jb, i64* first_zero, hipStream_t s)
{
    const i64 m = n - j0;
    const i64 below = m - jb;
    const unsigned g = (unsigned)(below > 0 ? (below + TS_NTH - 1) / TS_NTH : 1);
    hipLaunchKernelGGL(tslu_factor_kernel<T>, dim3(g), dim3(TS_NTH), 0, s, A, lda, j0, jb, m,
                       first_zero);
    return hip_status(hipGetLastError());
}

template int tslu_panel<float>(float*, i64, i64, i64, int, i64*, i64*, int*, hipStream_t);
template int tslu_factor<float>(float*, i64, i64, i64, int, i64*, hipStream_t);

} // namespace tmvx
