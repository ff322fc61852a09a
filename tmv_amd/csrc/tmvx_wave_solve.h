// One-wave solves with stored LU factors (n <= WS_NMAX), lane-distributed
// vectors: lane l holds entries l + 64 s, s < WS_PR.  Shared by the batched
// solve and the adjoint kernels.
//
// A x = b: permute by the interchange list, then unit-lower L and U by
// column sweeps (TMV LU_LDivEq, src/TMV_LUDiv.cpp:185-226, with the
// column-oriented substitution of src/TMV_TriDiv_V.cpp:147-218); a zero
// right-hand-side entry is skipped, as TMV does.
// A^T x = b: U^T then L^T as dot products over the contiguous columns of
// U and L (each a wave sum), then the interchanges in reverse order (TMV
// LU_RDivEq on b^T, src/TMV_LUDiv.cpp:369-410).
#pragma once
#include "tmvx_common.h"

namespace tmvx {

constexpr int WS_NMAX = 256;
constexpr int WS_PR = WS_NMAX / 64;

// Wave sum by DPP, result in every lane (quad swaps, mirrors, row
// broadcasts; rows outside a broadcast's mask add zero).
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_add(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xf, false);
    return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, 0xf,
                                                          false));
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    v = dpp_add<0xB1, 0xf>(v);      // quad_perm [1,0,3,2]
    v = dpp_add<0x4E, 0xf>(v);      // quad_perm [2,3,0,1]
    v = dpp_add<0x141, 0xf>(v);     // row_half_mirror
    v = dpp_add<0x140, 0xf>(v);     // row_mirror
    v = dpp_add<0x142, 0xa>(v);     // row_bcast:15
    v = dpp_add<0x143, 0xc>(v);     // row_bcast:31
    return readlane(v, 63);
}

template <typename T>
__device__ __forceinline__ void wave_load(T (&x)[WS_PR], const T* b, int n, int lane)
{
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        const int i = lane + 64 * s;
        x[s] = i < n ? b[i] : T(0);
    });
}

template <typename T>
__device__ __forceinline__ void wave_store(const T (&x)[WS_PR], T* b, int n, int lane)
{
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        const int i = lane + 64 * s;
        if (i < n) b[i] = x[s];
    });
}

// Entry k (uniform) of a lane-distributed vector, in every lane.
template <typename T>
__device__ __forceinline__ T wave_get(const T (&x)[WS_PR], int k)
{
    T v = x[0];
    static_for<1, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if ((k >> 6) == s) v = x[s];
    });
    return readlane(v, k & 63);
}

template <typename T>
__device__ __forceinline__ void wave_set(T (&x)[WS_PR], int k, T v, int lane)
{
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if (lane + 64 * s == k) x[s] = v;
    });
}

// Interchange entries j and p (uniform).
template <typename T>
__device__ __forceinline__ void wave_swap(T (&x)[WS_PR], int j, int p, int n, int lane)
{
    if (p == j || p < 0 || p >= n) return;
    const T xj = wave_get(x, j), xp = wave_get(x, p);
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        const int i = lane + 64 * s;
        if (i == j) x[s] = xp; else if (i == p) x[s] = xj;
    });
}

template <typename T>
__device__ __forceinline__ void wave_getrs(int trans, int n, const T* __restrict__ LU, i64 ld,
                                           const i64* __restrict__ piv, T (&x)[WS_PR], int lane)
{
    if (!trans) {
        for (int j = 0; j < n; ++j) wave_swap(x, j, (int)piv[j], n, lane);
        for (int k = 0; k < n; ++k) {                 // L, unit lower
            const T xk = wave_get(x, k);
            if (xk != T(0)) {
                static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    const int i = lane + 64 * s;
                    if (i > k && i < n) x[s] -= LU[i + (i64)k * ld] * xk;
                });
            }
        }
        for (int k = n - 1; k >= 0; --k) {            // U
            const T yk = wave_get(x, k);
            const T xk = yk == T(0) ? T(0) : yk / LU[k + (i64)k * ld];
            wave_set(x, k, xk, lane);
            if (xk != T(0)) {
                static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    const int i = lane + 64 * s;
                    if (i < k) x[s] -= LU[i + (i64)k * ld] * xk;
                });
            }
        }
    } else {
        for (int i = 0; i < n; ++i) {                 // U^T: lower, column i of U
            T part = T(0);
            static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                const int k = lane + 64 * s;
                if (k < i) part += LU[k + (i64)i * ld] * x[s];
            });
            const T gi = wave_get(x, i) - wave_sum(part);
            wave_set(x, i, gi == T(0) ? T(0) : gi / LU[i + (i64)i * ld], lane);
        }
        for (int i = n - 1; i >= 0; --i) {            // L^T: unit upper, column i of L
            T part = T(0);
            static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                const int k = lane + 64 * s;
                if (k > i && k < n) part += LU[k + (i64)i * ld] * x[s];
            });
            wave_set(x, i, wave_get(x, i) - wave_sum(part), lane);
        }
        for (int j = n - 1; j >= 0; --j) wave_swap(x, j, (int)piv[j], n, lane);
    }
}

} // namespace tmvx
