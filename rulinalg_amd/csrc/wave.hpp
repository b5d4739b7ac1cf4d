// Wave-level helpers for gfx950 (wave64): lane broadcasts and reductions over
// the wave built from DPP moves, which cost a few cycles each, where a
// shuffle through the LDS crossbar (ds_bpermute) costs about a hundred.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>
#include <utility>

namespace rlg {

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1 in order, so a
// loop body sees its index as a compile-time constant and registers indexed
// by it stay registers.
template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F &f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}


// Value of x in lane `l` (uniform result).
__device__ __forceinline__ int lane_get(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float lane_get(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_get(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
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
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
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

// ---------------------------------------------------------------- pivot search

// Candidate order: larger magnitude first; equal magnitudes go to the lower
// row index, which reproduces the reference's first-maximum scan.
template <class T>
__device__ __forceinline__ bool better(T v2, int i2, T v1, int i1) {
    // branch-free: the exec-mask branches of an if-chain cost more than the tests
    return (i2 != INT_MAX) & ((i1 == INT_MAX) | (v2 > v1) | (!(v1 > v2) & (i2 < i1)));
}

// The lane that holds the wave's first maximum under `better` (largest
// |value|, ties to the smallest row index; i == INT_MAX marks a lane without
// a candidate), or -1 if no lane has one; uniform.  One 64-bit maximum over
// the wave: in FP32 the key packs the value's bits (a non-negative float
// orders like its bits) above the complemented row index, so keys are
// distinct; in FP64 the key is the value's bits + 1 and a tie is settled by
// the smallest index among the tied lanes.  A NaN ranks like zero.
__device__ __forceinline__ int wave_first_max_lane(float v, int i) {
    unsigned long long key = 0;
    if (i != INT_MAX) {
        const unsigned b = (v == v) ? __float_as_uint(v) : 0u;
        key = ((unsigned long long)b << 32) | (0xFFFFFFFFu - (unsigned)i);
    }
    const unsigned long long m = wave_max_u64(key);
    if (m == 0) return -1;
    return __ffsll((long long)__ballot(key == m)) - 1;
}
__device__ __forceinline__ int wave_first_max_lane(double v, int i) {
    unsigned long long key = 0;
    if (i != INT_MAX) key = (v == v) ? (unsigned long long)__double_as_longlong(v) + 1 : 1;
    const unsigned long long m = wave_max_u64(key);
    if (m == 0) return -1;
    const unsigned long long tied = __ballot(key == m);
    if (__popcll(tied) == 1) return __ffsll((long long)tied) - 1;
    const int ii = key == m ? i : INT_MAX;
    const int mi = wave_min_i32(ii);
    return __ffsll((long long)__ballot(ii == mi)) - 1;
}

}  // namespace rlg
