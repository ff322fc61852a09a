// Pivot search helpers shared by the blocked and the batched LU kernels:
// TMV's underflow threshold and a DPP wave argmax with TMV's tie rule
// (largest |a|, lowest index on ties: FindMaxAbsElement's strict '>',
// src/TMV_Vector.cpp:473-504).
#pragma once
#include "tmvx_common.h"
#include <cfloat>

namespace tmvx {

template <typename T> __device__ __forceinline__ T tiny();
template <> __device__ __forceinline__ double tiny<double>() { return DBL_MIN; }
template <> __device__ __forceinline__ float tiny<float>() { return FLT_MIN; }

// Candidate (|value|, row); better = larger value, then smaller row.
template <typename T>
__device__ __forceinline__ bool better(T v1, int i1, T v0, int i0)
{ return v1 > v0 || (v1 == v0 && i1 < i0); }

// One DPP step of the wave argmax: combine (v, i) with the candidate held by
// the lane selected by CTRL (row_mask RM).  Rows outside RM keep theirs.
template <int CTRL, int RM, typename T>
__device__ __forceinline__ void dpp_argmax_step(T& v, int& i)
{
    T v2;
    if constexpr (sizeof(T) == 8) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, RM, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, RM, 0xf, false);
        v2 = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    } else {
        v2 = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                        CTRL, RM, 0xf, false));
    }
    const int i2 = __builtin_amdgcn_update_dpp(i, i, CTRL, RM, 0xf, false);
    if (better(v2, i2, v, i)) { v = v2; i = i2; }
}

// Wave argmax by DPP (quad swaps, half-row and row mirrors, row broadcasts);
// the result is returned to every lane from lane 63.
template <typename T>
__device__ __forceinline__ void wave_argmax(T& v, int& i)
{
    dpp_argmax_step<0xB1, 0xf>(v, i);    // quad_perm [1,0,3,2]
    dpp_argmax_step<0x4E, 0xf>(v, i);    // quad_perm [2,3,0,1]
    dpp_argmax_step<0x141, 0xf>(v, i);   // row_half_mirror
    dpp_argmax_step<0x140, 0xf>(v, i);   // row_mirror
    dpp_argmax_step<0x142, 0xa>(v, i);   // row_bcast:15
    dpp_argmax_step<0x143, 0xc>(v, i);   // row_bcast:31
    v = readlane(v, 63);
    i = __builtin_amdgcn_readlane(i, 63);
}


// Two-pass argmax with TMV's tie rule, cheaper on the dependency chain
// than the combined step above: the maximum value alone by a DPP max per
// stage, then a ballot of the lanes holding it; only an exact tie takes a
// second reduction, the minimum index among the tied lanes.  NaN values
// never win (fmax ignores them); with no comparable value the result is
// (-1, INT_MAX).  NL = 64 reduces the wave, NL = 8 lanes 0..7.
template <int CTRL, int RM, typename T>
__device__ __forceinline__ T dpp_max_step(T v)
{
    if constexpr (sizeof(T) == 8) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, RM, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, RM, 0xf, false);
        return fmax(v, __longlong_as_double(((long long)hi << 32) | (unsigned)lo));
    } else {
        return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                                   CTRL, RM, 0xf, false)));
    }
}

template <int CTRL, int RM>
__device__ __forceinline__ int dpp_min_step(int i)
{
    return min(i, __builtin_amdgcn_update_dpp(i, i, CTRL, RM, 0xf, false));
}

template <int NL, typename T>
__device__ __forceinline__ void lanes_argmax(T v, int i, T& bv, int& bi)
{
    static_assert(NL == 64 || NL == 8, "64 or 8 lanes");
    const int lane = threadIdx.x & 63;
    T m = v;
    m = dpp_max_step<0xB1, 0xf>(m);      // quad_perm [1,0,3,2]
    m = dpp_max_step<0x4E, 0xf>(m);      // quad_perm [2,3,0,1]
    m = dpp_max_step<0x141, 0xf>(m);     // row_half_mirror
    if constexpr (NL == 64) {
        m = dpp_max_step<0x140, 0xf>(m); // row_mirror
        m = dpp_max_step<0x142, 0xa>(m); // row_bcast:15
        m = dpp_max_step<0x143, 0xc>(m); // row_bcast:31
        m = readlane(m, 63);
    } else {
        m = readlane(m, 0);
    }
    const bool in = NL == 64 || lane < NL;
    const unsigned long long hit = __ballot(in && v == m);
    if (hit == 0ull) { bv = T(-1); bi = INT_MAX; return; }
    bv = m;
    if ((hit & (hit - 1)) == 0ull) {       // one lane holds the maximum
        bi = __builtin_amdgcn_readlane(i, (int)__builtin_ctzll(hit));
        return;
    }
    int x = (in && v == m) ? i : INT_MAX;  // an exact tie: lowest index
    x = dpp_min_step<0xB1, 0xf>(x);
    x = dpp_min_step<0x4E, 0xf>(x);
    x = dpp_min_step<0x141, 0xf>(x);
    if constexpr (NL == 64) {
        x = dpp_min_step<0x140, 0xf>(x);
        x = dpp_min_step<0x142, 0xa>(x);
        x = dpp_min_step<0x143, 0xc>(x);
        bi = __builtin_amdgcn_readlane(x, 63);
    } else {
        bi = __builtin_amdgcn_readlane(x, 0);
    }
}

} // namespace tmvx
