// GEMM tile configuration type and the per-configuration launcher
// declaration (defined in tmvx_gemm_kernel.h, instantiated per
// configuration in generated translation units).
#pragma once
#include "tmvx_common.h"

namespace tmvx {

template <typename T, int BM_, int BN_, int BK_, int WM_, int WN_, int MINW_>
struct GemmCfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_, MINW = MINW_;
    static constexpr int NT = 64 * WM * WN;          // threads
    static constexpr int V = Vec16<T>::n;            // elements per 16 B
    static constexpr int TM = BM / WM, TN = BN / WN; // wave tile
    static constexpr int FM = TM / 16, FN = TN / 16; // MFMA blocks per wave
    static constexpr int PADK = 16 / sizeof(T);      // pad of k-contiguous rows (16 B)
    static constexpr int PADX = 16 / sizeof(T);      // pad of k-major rows (16 B)
    static constexpr int NVA = BM * BK / V / NT;     // 16 B loads per thread
    static constexpr int NVB = BN * BK / V / NT;
    static_assert(BM % (16 * WM) == 0 && BN % (16 * WN) == 0, "wave tile");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");
    static_assert((BM * BK) % (V * NT) == 0 && (BN * BK) % (V * NT) == 0, "loads");
};

template <typename Cfg, typename T>
int gemm_launch_cfg(int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                    i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc,
                    hipStream_t s);

} // namespace tmvx
