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


} // namespace tmvx
