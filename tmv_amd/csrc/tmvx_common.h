// Shared definitions for the tmvx HIP kernels (gfx950 / CDNA4 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <type_traits>

#include "../../include/tmvx.h"

namespace tmvx {

typedef int64_t i64;

// MFMA 16x16x4 wrappers.  Operand layout (one value per lane, lane l,
// g = l/16):  a = A[i = l%16][k = g],  b = B[k = g][j = l%16].
// Result acc[r] = C[row(g, r)][j = l%16] with (measured on gfx950)
//   f64 16x16x4: row = g + 4r   (rows interleaved across the 4 registers)
//   f32 16x16x4: row = 4g + r   (4 consecutive rows per lane)
template <typename T> struct Mfma;

template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c)
    { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};

template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c)
    { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// 16-byte vector of T.
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); enum { n = 2 }; };
template <> struct Vec16<float>  { typedef float  type __attribute__((ext_vector_type(4))); enum { n = 4 }; };

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
// Keeps register-array indices static so arrays stay in VGPRs.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Broadcast lane `l`'s value of x to the whole wave (v_readlane).
__device__ __forceinline__ double readlane(double x, int l)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane(float x, int l)
{ return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }

inline int hip_status(hipError_t e) { return e == hipSuccess ? TMVX_OK : TMVX_ERR_HIP; }

#define TMVX_CHECK_LAUNCH()                                                  \
    do {                                                                     \
        hipError_t _e = hipGetLastError();                                   \
        if (_e != hipSuccess) return TMVX_ERR_HIP;                           \
    } while (0)

#define TMVX_TRY(expr)                                                       \
    do {                                                                     \
        int _rc = (expr);                                                    \
        if (_rc != TMVX_OK) return _rc;                                      \
    } while (0)

#define TMVX_HIP(expr)                                                       \
    do {                                                                     \
        hipError_t _e = (expr);                                              \
        if (_e != hipSuccess) return TMVX_ERR_HIP;                           \
    } while (0)

// Internal launchers shared across translation units.
template <typename T>
int gemm_launch(int a_rowmajor, int b_rowmajor, i64 m, i64 n, i64 k, T alpha,
                const T* A, i64 lda, const T* B, i64 ldb, T beta, T* C,
                i64 ldc, hipStream_t s);

// Complex C := (ar + i ai) op(A) op(B) + (br + i bi) C on interleaved
// complex storage of real type R (tmvx_complex.hip).
template <typename R>
int gemm_complex(int la, int lb, int conj_a, int conj_b, i64 m, i64 n, i64 k, R ar, R ai,
                 const R* A, i64 lda, const R* B, i64 ldb, R br, R bi, R* C, i64 ldc,
                 hipStream_t s);

} // namespace tmvx
