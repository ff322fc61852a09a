// GEMM entry points: configuration table, default choice per precision,
// and the C ABI (tmvx_gemm_f64 / tmvx_gemm_f32 and the tuning hooks).
#include "tmvx_gemm_decl.h"
#include "tmvx_gemm_cfgs.h"

namespace tmvx {

template <typename T> struct GemmTable;

#define TMVX_CFG_CASE(ID, BM, BN, BK, WM, WN, MW)                                   \
    case ID: return gemm_launch_cfg<GemmCfg<T, BM, BN, BK, WM, WN, MW>, T>(         \
                 ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
#define TMVX_CFG_NAME(ID, BM, BN, BK, WM, WN, MW)                                   \
    case ID: return #BM "x" #BN "x" #BK " waves " #WM "x" #WN " min/simd " #MW;
#define TMVX_CFG_COUNT(ID, BM, BN, BK, WM, WN, MW) +1

template <> struct GemmTable<double> {
    typedef double T;
    static constexpr int count = 0 TMVX_F64_CFGS(TMVX_CFG_COUNT);
    static int run(int cfg, int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                   i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc, hipStream_t s)
    {
        switch (cfg) { TMVX_F64_CFGS(TMVX_CFG_CASE) default: return TMVX_ERR_ARG; }
    }
    static const char* name(int cfg)
    { switch (cfg) { TMVX_F64_CFGS(TMVX_CFG_NAME) default: return nullptr; } }
};

template <> struct GemmTable<float> {
    typedef float T;
    static constexpr int count = 0 TMVX_F32_CFGS(TMVX_CFG_COUNT);
    static int run(int cfg, int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                   i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc, hipStream_t s)
    {
        switch (cfg) { TMVX_F32_CFGS(TMVX_CFG_CASE) default: return TMVX_ERR_ARG; }
    }
    static const char* name(int cfg)
    { switch (cfg) { TMVX_F32_CFGS(TMVX_CFG_NAME) default: return nullptr; } }
};

// Default configuration by shape (chosen from measurements on MI355X).
template <typename T> int default_cfg(int ta, int tb, i64 m, i64 n, i64 k);
// Short K (the LU updates: K = 64 inside an outer panel, K = 256 for the
// trailing matrix) favours small tiles: the C read-modify-write dominates
// and occupancy hides it (measured on MI355X at M = N = 4000..16000).
template <> int default_cfg<double>(int ta, int tb, i64 m, i64 n, i64 k)
{
    if (m * n < 1024 * 1024) return (m * n <= 256 * 256) ? 4 : 0;
    if (k <= 128) return 0;
    return (!ta && !tb) ? 1 : 0;
}
template <> int default_cfg<float>(int ta, int tb, i64 m, i64 n, i64 k)
{
    if (m * n < 1024 * 1024 || k <= 128) return 3;
    if (k <= 512) return 4;
    if (!ta && !tb) return 0;
    if (ta && !tb) return 2;
    return 1;
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
int gemm_run(int cfg, int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
             i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc, hipStream_t s)
{
    if (m < 0 || n < 0 || k < 0) return TMVX_ERR_ARG;
    if (m == 0 || n == 0) return TMVX_OK;
    if (ldc < m || lda < (ta ? k : m) || ldb < (tb ? n : k)) return TMVX_ERR_ARG;
    if (lda < 1 || ldb < 1) return TMVX_ERR_ARG;
    if (!C || (k > 0 && alpha != T(0) && (!A || !B))) return TMVX_ERR_ARG;
    if (k == 0 || alpha == T(0)) {
        if (beta == T(1)) return TMVX_OK;
        if (n > 65535) return TMVX_ERR_UNSUPPORTED;
        dim3 grid((unsigned)((m + 255) / 256), (unsigned)n);
        hipLaunchKernelGGL(scale_kernel<T>, grid, dim3(256), 0, s, m, n, beta, C, ldc);
        TMVX_CHECK_LAUNCH();
        return TMVX_OK;
    }
    if (cfg < 0) cfg = default_cfg<T>(ta, tb, m, n, k);
    return GemmTable<T>::run(cfg, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

template <typename T>
int gemm_launch(int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T* A,
                i64 lda, const T* B, i64 ldb, T beta, T* C, i64 ldc,
                hipStream_t s)
{
    return gemm_run<T>(-1, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

template int gemm_launch<double>(int, int, i64, i64, i64, double, const double*,
                                 i64, const double*, i64, double, double*, i64,
                                 hipStream_t);
template int gemm_launch<float>(int, int, i64, i64, i64, float, const float*,
                                i64, const float*, i64, float, float*, i64,
                                hipStream_t);

static bool layouts_ok(int la, int lb)
{
    return (la == TMVX_COLMAJOR || la == TMVX_ROWMAJOR) &&
           (lb == TMVX_COLMAJOR || lb == TMVX_ROWMAJOR);
}

// Row-major rows x cols (leading dimension ld) -> column-major copy with
// leading dimension rows: 64 x 64 tiles through LDS, reads along the rows
// and writes along the columns both coalesced.
template <typename T>
__global__ void __launch_bounds__(256)
to_colmajor_kernel(i64 rows, i64 cols, const T* __restrict__ X, i64 ld, T* __restrict__ Y)
{
    __shared__ T tile[64][65];
    const i64 r0 = (i64)blockIdx.y * 64, c0 = (i64)blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;     // 64 x 4
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = ty + 4 * q;
        if (r0 + r < rows && c0 + tx < cols) tile[r][tx] = X[(r0 + r) * ld + c0 + tx];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = ty + 4 * q;
        if (c0 + c < cols && r0 + tx < rows) Y[(r0 + tx) + (c0 + c) * rows] = tile[tx][c];
    }
}

template <typename T>
static int to_colmajor(i64 rows, i64 cols, const T* X, i64 ld, T* Y, hipStream_t s)
{
    const i64 gy = (rows + 63) / 64, gx = (cols + 63) / 64;
    if (gy > 65535 || gx > 0x7fffffff) return TMVX_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(to_colmajor_kernel<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s,
                       rows, cols, X, ld, Y);
    return hip_status(hipGetLastError());
}

// The public GEMM.  Large products with a row-major operand first copy that
// operand to column-major (TMV likewise copies operands into its blocked
// layout before its kernels, src/TMV_MultMM_Block.cpp:173-230): the
// column-major MFMA kernel runs at 78% of the FP64 peak at 8192^3 against
// 70% for the transposed-operand staging, and the copy costs one read and
// one write of the operand.  The arithmetic, and so the result, is the same.
template <typename T>
int gemm_public(int la, int lb, i64 m, i64 n, i64 k, T alpha, const T* A, i64 lda,
                const T* B, i64 ldb, T beta, T* C, i64 ldc, hipStream_t s)
{
    const bool big = (double)m * (double)n * (double)k >= 1073741824.0;   // 1024^3
    if (!big || (la == TMVX_COLMAJOR && lb == TMVX_COLMAJOR) || alpha == T(0) || k == 0 ||
        m <= 0 || n <= 0)
        return gemm_run<T>(-1, la, lb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
    if (ldc < m || lda < (la ? k : m) || ldb < (lb ? n : k) || !A || !B || !C)
        return TMVX_ERR_ARG;
    const size_t na = la ? (size_t)m * (size_t)k : 0, nb = lb ? (size_t)k * (size_t)n : 0;
    T* tmp = nullptr;
    if (hipMallocAsync((void**)&tmp, sizeof(T) * (na + nb), s) != hipSuccess) {
        (void)hipGetLastError();           // no room for the copy: stage in place
        return gemm_run<T>(-1, la, lb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
    }
    int rc = TMVX_OK;
    const T* Ac = A;
    const T* Bc = B;
    i64 ldac = lda, ldbc = ldb;
    if (la) { rc = to_colmajor<T>(m, k, A, lda, tmp, s); Ac = tmp; ldac = m; }
    if (rc == TMVX_OK && lb) { rc = to_colmajor<T>(k, n, B, ldb, tmp + na, s); Bc = tmp + na; ldbc = k; }
    if (rc == TMVX_OK)
        rc = gemm_run<T>(-1, TMVX_COLMAJOR, TMVX_COLMAJOR, m, n, k, alpha, Ac, ldac, Bc, ldbc,
                         beta, C, ldc, s);
    (void)hipFreeAsync(tmp, s);
    return rc;
}

} // namespace tmvx

extern "C" {

int tmvx_gemm_f64(int la, int lb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double* A, int64_t lda, const double* B, int64_t ldb,
                  double beta, double* C, int64_t ldc, void* stream)
{
    if (!tmvx::layouts_ok(la, lb)) return TMVX_ERR_ARG;
    return tmvx::gemm_public<double>(la, lb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                     (hipStream_t)stream);
}

int tmvx_gemm_f32(int la, int lb, int64_t m, int64_t n, int64_t k, float alpha,
                  const float* A, int64_t lda, const float* B, int64_t ldb,
                  float beta, float* C, int64_t ldc, void* stream)
{
    if (!tmvx::layouts_ok(la, lb)) return TMVX_ERR_ARG;
    return tmvx::gemm_public<float>(la, lb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                    (hipStream_t)stream);
}

int tmvx_gemm_config_count(int elem_bytes)
{
    return elem_bytes == 8 ? tmvx::GemmTable<double>::count
         : elem_bytes == 4 ? tmvx::GemmTable<float>::count : 0;
}

const char* tmvx_gemm_config_name(int elem_bytes, int cfg)
{
    const char* s = elem_bytes == 8 ? tmvx::GemmTable<double>::name(cfg)
                  : elem_bytes == 4 ? tmvx::GemmTable<float>::name(cfg) : nullptr;
    return s ? s : "";
}

int tmvx_gemm_cfg_f64(int cfg, int la, int lb, int64_t m, int64_t n, int64_t k, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb,
                      double beta, double* C, int64_t ldc, void* stream)
{
    if (!tmvx::layouts_ok(la, lb) || cfg < 0 || cfg >= tmvx::GemmTable<double>::count)
        return TMVX_ERR_ARG;
    return tmvx::gemm_run<double>(cfg, la, lb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                  (hipStream_t)stream);
}

int tmvx_gemm_cfg_f32(int cfg, int la, int lb, int64_t m, int64_t n, int64_t k, float alpha,
                      const float* A, int64_t lda, const float* B, int64_t ldb,
                      float beta, float* C, int64_t ldc, void* stream)
{
    if (!tmvx::layouts_ok(la, lb) || cfg < 0 || cfg >= tmvx::GemmTable<float>::count)
        return TMVX_ERR_ARG;
    return tmvx::gemm_run<float>(cfg, la, lb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                 (hipStream_t)stream);
}

} // extern "C"
