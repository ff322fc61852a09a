// MFMA GEMM kernel template for gfx950: C = alpha * op(A) * op(B) + beta * C.
//
// Counterpart of TMV's BlockMultMM + sse2_multmm_16_16_32 (src/TMV_MultMM_
// Block.cpp:173-440, src/TMV_MultMM_Kernel.h:2318-2512) and the simple
// CCC / CRC / RCC kernels (src/TMV_MultMM_{CCC,CRC,RCC}.cpp).  Instead of
// TMV's host-side packing into 16 x KB panels, each workgroup stages a
// BM x BK slice of A and a BK x BN slice of B in LDS (double buffered),
// loaded from HBM with 16-byte coalesced loads along whichever dimension is
// contiguous and stored to LDS in that same orientation (no transposes):
//   A column-major -> As[k][m] ("k-major"),  A row-major -> As[m][k];
//   B column-major -> Bs[n][k],              B row-major -> Bs[k][n].
// The MFMA k index is permuted inside each 16-wide k group: at substep s,
// lane group g = lane/16 supplies k = 4g + s for both operands, so a
// k-contiguous tile row is read as one 32-byte run per lane per 16 k.
//
// Workgroup = WM x WN waves; each wave owns a (BM/WM) x (BN/WN) block of C
// in MFMA accumulators (16x16 blocks of v_mfma_{f64,f32}_16x16x4).
#pragma once
#include "tmvx_gemm_decl.h"

namespace tmvx {

// XCD-aware tile order: workgroup w is dispatched to XCD w % 8; renumber so
// each XCD gets a contiguous run of tiles, then walk that run in groups of
// GROUP_M tile-rows so neighbouring tiles share A and B panels in one L2.
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn)
{
    constexpr int GROUP_M = 8;
    int w = blockIdx.x;
    const int nt = gridDim.x;
    if ((nt & 7) == 0) w = (w & 7) * (nt >> 3) + (w >> 3);
    const int group = GROUP_M * tiles_n;
    const int gid = w / group;
    const int first_m = gid * GROUP_M;
    const int gsz = min(tiles_m - first_m, GROUP_M);
    const int r = w - gid * group;
    tm = first_m + r % gsz;
    tn = r / gsz;
}

// One operand tile (rows x BK of op(X), "rows" = BM for A, BN for B) held in
// LDS.  KC (k contiguous in memory) selects the orientation.
template <typename T, int ROWS, int BK, bool KC>
struct TileSmem {
    static constexpr int PAD = 16 / sizeof(T);
    static constexpr int LD = KC ? (BK + PAD) : (ROWS + PAD);
    static constexpr int SIZE = KC ? ROWS * LD : BK * LD;   // elements per buffer
    // element (row x, k)
    static __device__ __forceinline__ int idx(int x, int k) { return KC ? x * LD + k : k * LD + x; }
};

template <typename Cfg, typename T, bool TA, bool TB>
__global__ void __launch_bounds__(Cfg::NT) __attribute__((amdgpu_waves_per_eu(Cfg::MINW)))
gemm_kernel(i64 M, i64 N, i64 K, T alpha, const T* __restrict__ A, i64 lda,
            const T* __restrict__ B, i64 ldb, T beta, T* __restrict__ C,
            i64 ldc, int vec_ok)
{
    constexpr int BM = Cfg::BM, BN = Cfg::BN, BK = Cfg::BK, WM = Cfg::WM;
    constexpr int NT = Cfg::NT, V = Cfg::V;
    typedef typename Vec16<T>::type VT;
    typedef typename Mfma<T>::v4 Acc;
    // A is k-contiguous in memory when row-major (TA); B when column-major.
    constexpr bool AKC = TA, BKC = !TB;
    typedef TileSmem<T, BM, BK, AKC> SA;
    typedef TileSmem<T, BN, BK, BKC> SB;

    __shared__ __attribute__((aligned(16))) T As[2][SA::SIZE];
    __shared__ __attribute__((aligned(16))) T Bs[2][SB::SIZE];

    const int tiles_m = (int)((M + BM - 1) / BM);
    const int tiles_n = (int)((N + BN - 1) / BN);
    int tm, tn;
    tile_coords(tiles_m, tiles_n, tm, tn);
    if (tn >= tiles_n) return;
    const i64 m0 = (i64)tm * BM, n0 = (i64)tn * BN;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int r16 = lane & 15, g = lane >> 4;

    Acc acc[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    VT ra[Cfg::NVA], rb[Cfg::NVB];
    const bool full_mn = (m0 + BM <= M) && (n0 + BN <= N);

    // Thread -> (row x, k) of its v-th 16-byte vector.  The vector runs
    // along k when the operand is k-contiguous, else along x.
    auto vcoord = [](int v, int rows, bool kc, int& x, int& k) {
        if (kc) { k = (v % (BK / V)) * V; x = v / (BK / V); }
        else    { x = (v % (rows / V)) * V; k = v / (rows / V); }
    };

    // ---- global -> registers -------------------------------------------
    auto load = [&](i64 k0) {
        const bool fast = vec_ok && full_mn && (k0 + BK <= K);
#pragma unroll
        for (int r = 0; r < Cfg::NVA; ++r) {
            int x, k;
            vcoord(t + r * NT, BM, AKC, x, k);
            if (fast) {
                const T* p = TA ? A + (m0 + x) * lda + (k0 + k) : A + (m0 + x) + (k0 + k) * lda;
                ra[r] = *reinterpret_cast<const VT*>(p);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const i64 gm = m0 + x + (AKC ? 0 : e), gk = k0 + k + (AKC ? e : 0);
                    ra[r][e] = (gm < M && gk < K) ? (TA ? A[gm * lda + gk] : A[gm + gk * lda]) : T(0);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < Cfg::NVB; ++r) {
            int x, k;
            vcoord(t + r * NT, BN, BKC, x, k);
            if (fast) {
                const T* p = TB ? B + (k0 + k) * ldb + (n0 + x) : B + (k0 + k) + (n0 + x) * ldb;
                rb[r] = *reinterpret_cast<const VT*>(p);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const i64 gn = n0 + x + (BKC ? 0 : e), gk = k0 + k + (BKC ? e : 0);
                    rb[r][e] = (gk < K && gn < N) ? (TB ? B[gk * ldb + gn] : B[gk + gn * ldb]) : T(0);
                }
            }
        }
    };

    // ---- registers -> LDS (same orientation, one 16-byte store each) ----
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < Cfg::NVA; ++r) {
            int x, k;
            vcoord(t + r * NT, BM, AKC, x, k);
            *reinterpret_cast<VT*>(&As[buf][SA::idx(x, k)]) = ra[r];
        }
#pragma unroll
        for (int r = 0; r < Cfg::NVB; ++r) {
            int x, k;
            vcoord(t + r * NT, BN, BKC, x, k);
            *reinterpret_cast<VT*>(&Bs[buf][SB::idx(x, k)]) = rb[r];
        }
    };

    // ---- MFMA over one LDS stage ---------------------------------------
    // Fragment of 4 k values (substeps s = 0..3) for lane group g:
    // k = kb + 4g + s.
    auto frag = [&](const T* S, int x, int kb, bool kc, int ld, T (&f)[4]) {
        if (kc) {
            const T* p = S + x * ld + kb + 4 * g;
#pragma unroll
            for (int s = 0; s < 4; ++s) f[s] = p[s];
        } else {
#pragma unroll
            for (int s = 0; s < 4; ++s) f[s] = S[(kb + 4 * g + s) * ld + x];
        }
    };
    auto compute = [&](int buf) {
        const int am = wm * Cfg::TM + r16;
        const int bn = wn * Cfg::TN + r16;
#pragma unroll
        for (int kb = 0; kb < BK; kb += 16) {
            T a[Cfg::FM][4], b[Cfg::FN][4];
#pragma unroll
            for (int i = 0; i < Cfg::FM; ++i) frag(As[buf], am + 16 * i, kb, AKC, SA::LD, a[i]);
#pragma unroll
            for (int j = 0; j < Cfg::FN; ++j) frag(Bs[buf], bn + 16 * j, kb, BKC, SB::LD, b[j]);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j)
                        acc[i][j] = Mfma<T>::mac(a[i][s], b[j][s], acc[i][j]);
        }
    };

    const i64 nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load(0);
        store(0);
        __syncthreads();
        for (i64 kt = 0; kt < nk; ++kt) {
            const int buf = (int)(kt & 1);
            if (kt + 1 < nk) load((kt + 1) * BK);
            compute(buf);
            if (kt + 1 < nk) store(buf ^ 1);
            __syncthreads();
        }
    }

    // ---- epilogue: acc[i][j][r] = C[row(g, r)][r16] of block (i, j) -----
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i) {
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) {
            const i64 col = n0 + wn * Cfg::TN + 16 * j + r16;
            const i64 rowb = m0 + wm * Cfg::TM + 16 * i;
            if (col >= N) continue;
            T* cp = C + col * ldc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 row = rowb + Mfma<T>::row(g, r);
                if (row < M) {
                    const T v = alpha * acc[i][j][r];
                    cp[row] = (beta == T(0)) ? v : v + beta * cp[row];
                }
            }
        }
    }
}

template <typename Cfg, typename T>
int gemm_launch_cfg(int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                    i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc,
                    hipStream_t s)
{
    const i64 tiles = ((m + Cfg::BM - 1) / Cfg::BM) * ((n + Cfg::BN - 1) / Cfg::BN);
    if (tiles > 0x7fffffff) return TMVX_ERR_UNSUPPORTED;
    const int V = Cfg::V;
    const int vec_ok =
        ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % 16 == 0) &&
        (lda % V == 0) && (ldb % V == 0);
    dim3 grid((unsigned)tiles), block(Cfg::NT);
#define TMVX_GEMM_GO(TA_, TB_)                                                 \
    hipLaunchKernelGGL((gemm_kernel<Cfg, T, TA_, TB_>), grid, block, 0, s, m,  \
                       n, k, alpha, A, lda, B, ldb, beta, C, ldc, vec_ok)
    if (!ta && !tb) TMVX_GEMM_GO(false, false);
    else if (!ta && tb) TMVX_GEMM_GO(false, true);
    else if (ta && !tb) TMVX_GEMM_GO(true, false);
    else TMVX_GEMM_GO(true, true);
#undef TMVX_GEMM_GO
    TMVX_CHECK_LAUNCH();
    return TMVX_OK;
}

} // namespace tmvx
