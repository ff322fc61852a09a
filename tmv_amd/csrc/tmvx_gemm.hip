// MFMA GEMM for gfx950: C = alpha * op(A) * op(B) + beta * C.
//
// Counterpart of TMV's BlockMultMM + sse2_multmm_16_16_32 (src/TMV_MultMM_
// Block.cpp:173-440, src/TMV_MultMM_Kernel.h:2318-2512) and of the simple
// CCC / CRC / RCC kernels (src/TMV_MultMM_{CCC,CRC,RCC}.cpp): one kernel
// template covers every operand layout.  Instead of TMV's host-side packing
// into 16xKB panels, each workgroup stages a BM x BK slice of A and a BK x BN
// slice of B through LDS (double buffered, k-major so both MFMA operands
// are read with one ds_read per lane), loading them from HBM with 16-byte
// coalesced loads along whichever dimension is contiguous in memory.
//
// Workgroup = WM x WN waves; each wave owns a (BM/WM) x (BN/WN) block of C,
// kept in MFMA accumulators (16x16 blocks of v_mfma_{f64,f32}_16x16x4).
#include "tmvx_common.h"

namespace tmvx {

template <typename T, int BM, int BN, int BK, int WM, int WN>
struct GemmCfg {
    static constexpr int NT = 64 * WM * WN;          // threads
    static constexpr int V = Vec16<T>::n;            // elements per 16 B
    static constexpr int TM = BM / WM, TN = BN / WN; // wave tile
    static constexpr int FM = TM / 16, FN = TN / 16; // MFMA blocks per wave
    static constexpr int PADA = 16 / sizeof(T) * 2;  // LDS row padding
    static constexpr int PADB = 16 / sizeof(T) * 2;
    static constexpr int LDA_S = BM + PADA, LDB_S = BN + PADB;
    static constexpr int NVA = BM * BK / V / NT;     // 16 B loads per thread
    static constexpr int NVB = BN * BK / V / NT;
    static_assert(BM % (16 * WM) == 0 && BN % (16 * WN) == 0, "wave tile");
    static_assert(BK % 4 == 0, "BK");
    static_assert((BM * BK) % (V * NT) == 0 && (BN * BK) % (V * NT) == 0, "loads");
};

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

template <typename T, int BM, int BN, int BK, int WM, int WN, bool TA, bool TB>
__global__ void __launch_bounds__(64 * WM * WN)
gemm_kernel(i64 M, i64 N, i64 K, T alpha, const T* __restrict__ A, i64 lda,
            const T* __restrict__ B, i64 ldb, T beta, T* __restrict__ C,
            i64 ldc, int vec_ok)
{
    typedef GemmCfg<T, BM, BN, BK, WM, WN> Cfg;
    typedef typename Vec16<T>::type VT;
    typedef typename Mfma<T>::v4 Acc;
    constexpr int NT = Cfg::NT, V = Cfg::V;

    __shared__ __attribute__((aligned(16))) T As[2][BK][Cfg::LDA_S];
    __shared__ __attribute__((aligned(16))) T Bs[2][BK][Cfg::LDB_S];

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

    // ---- global -> registers -------------------------------------------
    auto load = [&](i64 k0) {
        const bool fast = vec_ok && full_mn && (k0 + BK <= K);
#pragma unroll
        for (int r = 0; r < Cfg::NVA; ++r) {
            const int v = t + r * NT;
            int m, k;
            if (!TA) { m = (v % (BM / V)) * V; k = v / (BM / V); }
            else     { k = (v % (BK / V)) * V; m = v / (BK / V); }
            if (fast) {
                const T* p = TA ? A + (m0 + m) * lda + (k0 + k)
                                : A + (m0 + m) + (k0 + k) * lda;
                ra[r] = *reinterpret_cast<const VT*>(p);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const i64 gm = m0 + m + (TA ? 0 : e), gk = k0 + k + (TA ? e : 0);
                    ra[r][e] = (gm < M && gk < K)
                        ? (TA ? A[gm * lda + gk] : A[gm + gk * lda]) : T(0);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < Cfg::NVB; ++r) {
            const int v = t + r * NT;
            int n, k;
            if (!TB) { k = (v % (BK / V)) * V; n = v / (BK / V); }
            else     { n = (v % (BN / V)) * V; k = v / (BN / V); }
            if (fast) {
                const T* p = TB ? B + (k0 + k) * ldb + (n0 + n)
                                : B + (k0 + k) + (n0 + n) * ldb;
                rb[r] = *reinterpret_cast<const VT*>(p);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const i64 gk = k0 + k + (TB ? 0 : e), gn = n0 + n + (TB ? e : 0);
                    rb[r][e] = (gk < K && gn < N)
                        ? (TB ? B[gk * ldb + gn] : B[gk + gn * ldb]) : T(0);
                }
            }
        }
    };

    // ---- registers -> LDS (k-major: As[k][m], Bs[k][n]) ----------------
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < Cfg::NVA; ++r) {
            const int v = t + r * NT;
            if (!TA) {
                const int m = (v % (BM / V)) * V, k = v / (BM / V);
                *reinterpret_cast<VT*>(&As[buf][k][m]) = ra[r];
            } else {
                const int k = (v % (BK / V)) * V, m = v / (BK / V);
#pragma unroll
                for (int e = 0; e < V; ++e) As[buf][k + e][m] = ra[r][e];
            }
        }
#pragma unroll
        for (int r = 0; r < Cfg::NVB; ++r) {
            const int v = t + r * NT;
            if (!TB) {
                const int k = (v % (BK / V)) * V, n = v / (BK / V);
#pragma unroll
                for (int e = 0; e < V; ++e) Bs[buf][k + e][n] = rb[r][e];
            } else {
                const int n = (v % (BN / V)) * V, k = v / (BN / V);
                *reinterpret_cast<VT*>(&Bs[buf][k][n]) = rb[r];
            }
        }
    };

    // ---- MFMA over one LDS stage ---------------------------------------
    auto compute = [&](int buf) {
        const int am = wm * Cfg::TM + r16;
        const int bn = wn * Cfg::TN + r16;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[Cfg::FM], b[Cfg::FN];
#pragma unroll
            for (int i = 0; i < Cfg::FM; ++i) a[i] = As[buf][kk + g][am + 16 * i];
#pragma unroll
            for (int j = 0; j < Cfg::FN; ++j) b[j] = Bs[buf][kk + g][bn + 16 * j];
#pragma unroll
            for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                for (int j = 0; j < Cfg::FN; ++j)
                    acc[i][j] = Mfma<T>::mac(a[i], b[j], acc[i][j]);
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

template <typename T, int BM, int BN, int BK, int WM, int WN>
static int launch_cfg(int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                      i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc,
                      hipStream_t s)
{
    typedef GemmCfg<T, BM, BN, BK, WM, WN> Cfg;
    const i64 tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    if (tiles > 0x7fffffff) return TMVX_ERR_UNSUPPORTED;
    const int V = Cfg::V;
    const int vec_ok =
        ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % 16 == 0) &&
        (lda % V == 0) && (ldb % V == 0);
    dim3 grid((unsigned)tiles), block(Cfg::NT);
#define TMVX_GEMM_GO(TA_, TB_)                                                 \
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WM, WN, TA_, TB_>), grid,   \
                       block, 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C,   \
                       ldc, vec_ok)
    if (!ta && !tb) TMVX_GEMM_GO(false, false);
    else if (!ta && tb) TMVX_GEMM_GO(false, true);
    else if (ta && !tb) TMVX_GEMM_GO(true, false);
    else TMVX_GEMM_GO(true, true);
#undef TMVX_GEMM_GO
    TMVX_CHECK_LAUNCH();
    return TMVX_OK;
}

// Scale C by beta (k == 0 or alpha == 0): C = beta * C, or 0 when beta == 0.
template <typename T>
__global__ void scale_kernel(i64 m, i64 n, T beta, T* C, i64 ldc)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 j = blockIdx.y;
    if (i < m && j < n) C[i + j * ldc] = beta == T(0) ? T(0) : beta * C[i + j * ldc];
}

template <typename T>
int gemm_launch(int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc,
                hipStream_t s)
{
    if (m < 0 || n < 0 || k < 0) return TMVX_ERR_ARG;
    if (m == 0 || n == 0) return TMVX_OK;
    if (ldc < m || lda < (ta ? k : m) || ldb < (tb ? n : k)) return TMVX_ERR_ARG;
    if (!C || (k > 0 && alpha != T(0) && (!A || !B))) return TMVX_ERR_ARG;
    if (k == 0 || alpha == T(0)) {
        if (beta == T(1)) return TMVX_OK;
        if (n > 65535) return TMVX_ERR_UNSUPPORTED;
        dim3 grid((unsigned)((m + 255) / 256), (unsigned)n);
        hipLaunchKernelGGL(scale_kernel<T>, grid, dim3(256), 0, s, m, n, beta, C, ldc);
        TMVX_CHECK_LAUNCH();
        return TMVX_OK;
    }
    if constexpr (sizeof(T) == 8)
        return launch_cfg<T, 128, 128, 16, 2, 2>(ta, tb, m, n, k, alpha, A, lda,
                                                 B, ldb, beta, C, ldc, s);
    else
        return launch_cfg<T, 128, 128, 32, 2, 2>(ta, tb, m, n, k, alpha, A, lda, B,
                                                 ldb, beta, C, ldc, s);
}

template int gemm_launch<double>(int, int, i64, i64, i64, double, const double*,
                                 i64, const double*, i64, double, double*, i64,
                                 hipStream_t);
template int gemm_launch<float>(int, int, i64, i64, i64, float, const float*,
                                i64, const float*, i64, float, float*, i64,
                                hipStream_t);

} // namespace tmvx

extern "C" {

int tmvx_gemm_f64(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double* A, int64_t lda, const double* B, int64_t ldb,
                  double beta, double* C, int64_t ldc, void* stream)
{
    if ((la != TMVX_COLMAJOR && la != TMVX_ROWMAJOR) ||
        (lb != TMVX_COLMAJOR && lb != TMVX_ROWMAJOR)) return TMVX_ERR_ARG;
    return tmvx::gemm_launch<double>(la, lb, m, n, k, alpha, A, lda, B, ldb,
                                     beta, C, ldc, (hipStream_t)stream);
}

int tmvx_gemm_f32(int la, int lb, int64_t m, int64_t n, int64_t k, float alpha,
                  const float* A, int64_t lda, const float* B, int64_t ldb,
                  float beta, float* C, int64_t ldc, void* stream)
{
    if ((la != TMVX_COLMAJOR && la != TMVX_ROWMAJOR) ||
        (lb != TMVX_COLMAJOR && lb != TMVX_ROWMAJOR)) return TMVX_ERR_ARG;
    return tmvx::gemm_launch<float>(la, lb, m, n, k, alpha, A, lda, B, ldb,
                                    beta, C, ldc, (hipStream_t)stream);
}

} // extern "C"
