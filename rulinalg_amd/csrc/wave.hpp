// Wave-level helpers for gfx950 (wave64): lane broadcasts and reductions over
// the wave built from DPP moves, which cost a few cycles each, where a
// shuffle through the LDS crossbar (ds_bpermute) costs about a hundred.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace rlg {

// Value of x in lane `l` (uniform result).
__device__ __forceinline__ int lane_get(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float lane_get(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_get(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// DPP lane permutations within rows of 16 lanes.
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2, 3, 0, 1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7 - i within each 8
constexpr int DPP_MIRROR = 0x140;      // lane i <-> 15 - i within each 16

template <int CTRL> __device__ __forceinline__ int dpp(int x) {
    return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ float dpp(float x) {
    return __int_as_float(dpp<CTRL>(__float_as_int(x)));
}
template <int CTRL> __device__ __forceinline__ double dpp(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp<CTRL>((int)b), hi = dpp<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ unsigned long long lane_get(unsigned long long x, int l) {
    const int lo = __builtin_amdgcn_readlane((int)x, l);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
template <int CTRL> __device__ __forceinline__ unsigned long long dpp(unsigned long long x) {
    const int lo = dpp<CTRL>((int)x), hi = dpp<CTRL>((int)(x >> 32));
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Maximum of a 64-bit key over the wave (uniform result): two quad swaps
// and two row mirrors, then the four rows' maxima from lanes 0, 16, 32, 48.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
    auto mx = [](unsigned long long a, unsigned long long b) { return a > b ? a : b; };
    k = mx(k, dpp<DPP_XOR1>(k));
    k = mx(k, dpp<DPP_XOR2>(k));
    k = mx(k, dpp<DPP_HALF_MIRROR>(k));
    k = mx(k, dpp<DPP_MIRROR>(k));
    return mx(mx(lane_get(k, 0), lane_get(k, 16)), mx(lane_get(k, 32), lane_get(k, 48)));
}

// Sum over each row of 16 lanes, in every lane of the row: four rotations
// within the row (row_ror 8, 4, 2, 1), each a DPP move.
template <int N> constexpr int DPP_ROR = 0x120 + N;
template <class T> __device__ __forceinline__ T row16_sum(T a) {
    a += dpp<DPP_ROR<8>>(a);
    a += dpp<DPP_ROR<4>>(a);
    a += dpp<DPP_ROR<2>>(a);
    a += dpp<DPP_ROR<1>>(a);
    return a;
}

// Minimum of an int over the wave (uniform result), the same pattern.
__device__ __forceinline__ int wave_min_i32(int k) {
    k = min(k, dpp<DPP_XOR1>(k));
    k = min(k, dpp<DPP_XOR2>(k));
    k = min(k, dpp<DPP_HALF_MIRROR>(k));
    k = min(k, dpp<DPP_MIRROR>(k));
    return min(min(lane_get(k, 0), lane_get(k, 16)), min(lane_get(k, 32), lane_get(k, 48)));
}

}  // namespace rlg
