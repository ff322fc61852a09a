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

// Column sweeps with prefetch: BODY runs for the columns k of LU in sweep
// order (forward 0..n-1 or backward n-1..0) with col[s] = LU(lane + 64 s,
// k) (rows past n read as 0).  Eight columns are in flight: the loop is
// unrolled by eight with one named register vector per slot, and a slot's
// next column (eight steps ahead) is loaded right after the slot is
// consumed, so a step's loads are issued well before the dependency chain
// reaches them.  (Named vectors, not an array: an array of buffers used in
// turn ended up in scratch memory.)
template <typename T> struct WsCol { typedef T type __attribute__((ext_vector_type(WS_PR))); };

template <typename T>
__device__ __forceinline__ typename WsCol<T>::type ws_load_col(const T* __restrict__ LU, i64 ld,
                                                                int n, int q, bool backward,
                                                                int lane)
{
    typename WsCol<T>::type c;
    const int k = q < n ? (backward ? n - 1 - q : q) : 0;
    const T* cp = LU + (i64)k * ld;
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        const int i = lane + 64 * s;
        c[s] = (q < n && i < n) ? cp[i] : T(0);
    });
    return c;
}

// Entry k (uniform) of a column vector, in every lane.
template <typename V>
__device__ __forceinline__ auto wave_get_v(const V& x, int k)
{
    auto v = x[0];
    static_for<1, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if ((k >> 6) == s) v = x[s];
    });
    return readlane(v, k & 63);
}

#define TMVX_WS_STEP(SLOT, Q, D, BACKWARD, ...)                                              \
    if ((Q) < n) {                                                                         \
        const int k = (BACKWARD) ? n - 1 - (Q) : (Q);                                      \
        const typename WsCol<T>::type col = SLOT;                                          \
        __VA_ARGS__                                                                        \
    }                                                                                      \
    SLOT = ws_load_col(LU, ld, n, (Q) + (D), (BACKWARD), lane);
#define TMVX_WS_SWEEP(BACKWARD, ...)                                                        \
    do {                                                                                    \
        typename WsCol<T>::type p0_ = ws_load_col(LU, ld, n, 0, (BACKWARD), lane);          \
        typename WsCol<T>::type p1_ = ws_load_col(LU, ld, n, 1, (BACKWARD), lane);          \
        typename WsCol<T>::type p2_ = ws_load_col(LU, ld, n, 2, (BACKWARD), lane);          \
        typename WsCol<T>::type p3_ = ws_load_col(LU, ld, n, 3, (BACKWARD), lane);          \
        typename WsCol<T>::type p4_ = ws_load_col(LU, ld, n, 4, (BACKWARD), lane);          \
        typename WsCol<T>::type p5_ = ws_load_col(LU, ld, n, 5, (BACKWARD), lane);          \
        typename WsCol<T>::type p6_ = ws_load_col(LU, ld, n, 6, (BACKWARD), lane);          \
        typename WsCol<T>::type p7_ = ws_load_col(LU, ld, n, 7, (BACKWARD), lane);          \
        for (int q0_ = 0; q0_ < n; q0_ += 8) {                                              \
            TMVX_WS_STEP(p0_, q0_, 8, (BACKWARD), __VA_ARGS__)                              \
            TMVX_WS_STEP(p1_, q0_ + 1, 8, (BACKWARD), __VA_ARGS__)                          \
            TMVX_WS_STEP(p2_, q0_ + 2, 8, (BACKWARD), __VA_ARGS__)                          \
            TMVX_WS_STEP(p3_, q0_ + 3, 8, (BACKWARD), __VA_ARGS__)                          \
            TMVX_WS_STEP(p4_, q0_ + 4, 8, (BACKWARD), __VA_ARGS__)                          \
            TMVX_WS_STEP(p5_, q0_ + 5, 8, (BACKWARD), __VA_ARGS__)                          \
            TMVX_WS_STEP(p6_, q0_ + 6, 8, (BACKWARD), __VA_ARGS__)                          \
            TMVX_WS_STEP(p7_, q0_ + 7, 8, (BACKWARD), __VA_ARGS__)                          \
        }                                                                                   \
    } while (0)
// the same with sixteen columns in flight (for callers with registers to
// spare: the latency of a column load is then spread over sixteen steps)
#define TMVX_WS_SWEEP16(BACKWARD, ...)                                                      \
    do {                                                                                    \
        typename WsCol<T>::type p0_ = ws_load_col(LU, ld, n, 0, (BACKWARD), lane);          \
        typename WsCol<T>::type p1_ = ws_load_col(LU, ld, n, 1, (BACKWARD), lane);          \
        typename WsCol<T>::type p2_ = ws_load_col(LU, ld, n, 2, (BACKWARD), lane);          \
        typename WsCol<T>::type p3_ = ws_load_col(LU, ld, n, 3, (BACKWARD), lane);          \
        typename WsCol<T>::type p4_ = ws_load_col(LU, ld, n, 4, (BACKWARD), lane);          \
        typename WsCol<T>::type p5_ = ws_load_col(LU, ld, n, 5, (BACKWARD), lane);          \
        typename WsCol<T>::type p6_ = ws_load_col(LU, ld, n, 6, (BACKWARD), lane);          \
        typename WsCol<T>::type p7_ = ws_load_col(LU, ld, n, 7, (BACKWARD), lane);          \
        typename WsCol<T>::type p8_ = ws_load_col(LU, ld, n, 8, (BACKWARD), lane);          \
        typename WsCol<T>::type p9_ = ws_load_col(LU, ld, n, 9, (BACKWARD), lane);          \
        typename WsCol<T>::type p10_ = ws_load_col(LU, ld, n, 10, (BACKWARD), lane);        \
        typename WsCol<T>::type p11_ = ws_load_col(LU, ld, n, 11, (BACKWARD), lane);        \
        typename WsCol<T>::type p12_ = ws_load_col(LU, ld, n, 12, (BACKWARD), lane);        \
        typename WsCol<T>::type p13_ = ws_load_col(LU, ld, n, 13, (BACKWARD), lane);        \
        typename WsCol<T>::type p14_ = ws_load_col(LU, ld, n, 14, (BACKWARD), lane);        \
        typename WsCol<T>::type p15_ = ws_load_col(LU, ld, n, 15, (BACKWARD), lane);        \
        for (int q0_ = 0; q0_ < n; q0_ += 16) {                                             \
            TMVX_WS_STEP(p0_, q0_, 16, (BACKWARD), __VA_ARGS__)                             \
            TMVX_WS_STEP(p1_, q0_ + 1, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p2_, q0_ + 2, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p3_, q0_ + 3, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p4_, q0_ + 4, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p5_, q0_ + 5, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p6_, q0_ + 6, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p7_, q0_ + 7, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p8_, q0_ + 8, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p9_, q0_ + 9, 16, (BACKWARD), __VA_ARGS__)                         \
            TMVX_WS_STEP(p10_, q0_ + 10, 16, (BACKWARD), __VA_ARGS__)                       \
            TMVX_WS_STEP(p11_, q0_ + 11, 16, (BACKWARD), __VA_ARGS__)                       \
            TMVX_WS_STEP(p12_, q0_ + 12, 16, (BACKWARD), __VA_ARGS__)                       \
            TMVX_WS_STEP(p13_, q0_ + 13, 16, (BACKWARD), __VA_ARGS__)                       \
            TMVX_WS_STEP(p14_, q0_ + 14, 16, (BACKWARD), __VA_ARGS__)                       \
            TMVX_WS_STEP(p15_, q0_ + 15, 16, (BACKWARD), __VA_ARGS__)                       \
        }                                                                                   \
    } while (0)

// Interchange list entry j (uniform) from its lane-distributed copy.
__device__ __forceinline__ int wave_get_int(const int (&v)[WS_PR], int j)
{
    int r = v[0];
    static_for<1, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if ((j >> 6) == s) r = v[s];
    });
    return __builtin_amdgcn_readlane(r, j & 63);
}

template <typename T, bool DEEP = false>
__device__ __forceinline__ void wave_getrs(int trans, int n, const T* __restrict__ LU, i64 ld,
                                           const i64* __restrict__ piv, T (&x)[WS_PR], int lane)
{
    // the interchange list in registers (one coalesced load, no load on the
    // chain of swaps)
    int pv[WS_PR];
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        const int i = lane + 64 * s;
        pv[s] = i < n ? (int)piv[i] : i;
    });
    if (!trans) {
        for (int j = 0; j < n; ++j) wave_swap(x, j, wave_get_int(pv, j), n, lane);
        TMVX_WS_SWEEP(false, {                                  // L, unit lower
            const T xk = wave_get(x, k);
            if (xk != T(0)) {
                static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    const int i = lane + 64 * s;
                    if (i > k && i < n) x[s] -= col[s] * xk;
                });
            }
        });
        TMVX_WS_SWEEP(true, {                                   // U
            const T yk = wave_get(x, k);
            const T xk = yk == T(0) ? T(0) : yk / wave_get_v(col, k);
            wave_set(x, k, xk, lane);
            if (xk != T(0)) {
                static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    const int i = lane + 64 * s;
                    if (i < k) x[s] -= col[s] * xk;
                });
            }
        });
    } else {
        // U^T: column k of U, then L^T: column k of L
#define TMVX_WS_UT_BODY                                                                    \
    T part = T(0);                                                                         \
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {                     \
        constexpr int s = decltype(sc)::value;                                             \
        const int r = lane + 64 * s;                                                       \
        if (r < k) part += col[s] * x[s];                                                  \
    });                                                                                    \
    const T gk = wave_get(x, k) - wave_sum(part);                                          \
    wave_set(x, k, gk == T(0) ? T(0) : gk / wave_get_v(col, k), lane);
#define TMVX_WS_LT_BODY                                                                    \
    T part = T(0);                                                                         \
    static_for<0, WS_PR>([&](auto sc) __attribute__((always_inline)) {                     \
        constexpr int s = decltype(sc)::value;                                             \
        const int r = lane + 64 * s;                                                       \
        if (r > k && r < n) part += col[s] * x[s];                                         \
    });                                                                                    \
    wave_set(x, k, wave_get(x, k) - wave_sum(part), lane);
        if constexpr (DEEP) {
            TMVX_WS_SWEEP16(false, { TMVX_WS_UT_BODY });
            TMVX_WS_SWEEP16(true, { TMVX_WS_LT_BODY });
        } else {
            TMVX_WS_SWEEP(false, { TMVX_WS_UT_BODY });
            TMVX_WS_SWEEP(true, { TMVX_WS_LT_BODY });
        }
#undef TMVX_WS_UT_BODY
#undef TMVX_WS_LT_BODY
        for (int j = n - 1; j >= 0; --j) wave_swap(x, j, wave_get_int(pv, j), n, lane);
    }
}

#undef TMVX_WS_SWEEP
#undef TMVX_WS_SWEEP16
#undef TMVX_WS_STEP

} // namespace tmvx
