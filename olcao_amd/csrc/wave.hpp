// Cross-lane operations of one 64-lane wave without LDS round trips.  A
// broadcast from a wave-uniform lane is a readlane; a wave sum (or maximum)
// is four DPP steps inside each row of 16 lanes (swap neighbours, swap
// pairs, mirror the halves, mirror the row) followed by the combination of
// the four row results read back by readlane, so that every lane holds the
// same value.  All lanes of the wave must be active.
#pragma once

#include <hip/hip_runtime.h>

namespace olcao {

__device__ __forceinline__ double rlane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int kDppQuadSwap1 = 0xb1;      // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4e;      // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;    // row_half_mirror
constexpr int kDppMirror = 0x140;        // row_mirror

__device__ __forceinline__ double wave_sum(double a)
{
    a += dpp<kDppQuadSwap1>(a);
    a += dpp<kDppQuadSwap2>(a);
    a += dpp<kDppHalfMirror>(a);
    a += dpp<kDppMirror>(a);
    return (rlane(a, 0) + rlane(a, 16)) + (rlane(a, 32) + rlane(a, 48));
}

__device__ __forceinline__ double wave_max(double a)
{
    a = fmax(a, dpp<kDppQuadSwap1>(a));
    a = fmax(a, dpp<kDppQuadSwap2>(a));
    a = fmax(a, dpp<kDppHalfMirror>(a));
    a = fmax(a, dpp<kDppMirror>(a));
    return fmax(fmax(rlane(a, 0), rlane(a, 16)), fmax(rlane(a, 32), rlane(a, 48)));
}

}  // namespace olcao
