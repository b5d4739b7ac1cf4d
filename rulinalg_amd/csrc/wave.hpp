// Wave-level helpers for gfx950 (wave64): lane broadcasts and a first-max
// (argmax) reduction over the wave built from DPP moves, which cost a few
// cycles each, where a shuffle through the LDS crossbar (ds_bpermute) costs
// about a hundred.
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

// Argmax over the wave of (v, i) under the strict order `Better`, carrying a
// payload; every lane ends with the winner.  The order is total, so any
// pairing of lanes gives the same result: two quad swaps and two mirrors
// leave each row of 16 lanes with its best, and the four rows are combined
// from lanes 0, 16, 32 and 48.
template <class T, class Better>
__device__ __forceinline__ void wave_argmax_dpp(T &v, int &i, int &pay, Better better) {
    auto step = [&](auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        const T ov = dpp<C>(v);
        const int oi = dpp<C>(i), op = dpp<C>(pay);
        if (better(ov, oi, v, i)) { v = ov; i = oi; pay = op; }
    };
    step(std::integral_constant<int, DPP_XOR1>{});
    step(std::integral_constant<int, DPP_XOR2>{});
    step(std::integral_constant<int, DPP_HALF_MIRROR>{});
    step(std::integral_constant<int, DPP_MIRROR>{});
    T bv = lane_get(v, 0);
    int bi = lane_get(i, 0), bp = lane_get(pay, 0);
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        const T qv = lane_get(v, 16 * q);
        const int qi = lane_get(i, 16 * q), qp = lane_get(pay, 16 * q);
        if (better(qv, qi, bv, bi)) { bv = qv; bi = qi; bp = qp; }
    }
    v = bv;
    i = bi;
    pay = bp;
}

}  // namespace rlg
