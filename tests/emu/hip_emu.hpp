// CPU emulation of the device facilities the batched kernels use (test
// infrastructure: built only into tests/emu's checker, never into librlgpu).
//
// A launch runs every lane of one workgroup as a fiber on one host thread
// (emu_runtime.cpp).  A lane runs until it reaches a synchronization point:
//   * a workgroup barrier, released once every live lane of the workgroup has
//     reached it from the same call site;
//   * a wave collective (lane read, first-lane read, DPP move, ballot, MFMA),
//     computed once every live lane of the wave has reached it from the same
//     call site with the same uniform operands (lane index, DPP control).
// Anything else (lanes of one wave or workgroup waiting at different sites,
// a non-uniform lane index, a lane that exits while others of its wave
// still compute) is reported as an error and ends the run.  Between
// synchronization points lanes run one after another, in an order the
// driver chooses, so that results which depend on that order expose a race.
#pragma once

#include <math.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

namespace emu {

struct Dim3 {
    unsigned x = 0, y = 0, z = 0;
};

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// The running lane's indices (set by the scheduler before it resumes a lane).
extern Dim3 g_thread, g_block, g_block_dim, g_grid_dim;

void syncthreads();
int readlane(int v, int lane);
int readfirstlane(int v);
int update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl);
unsigned long long ballot(bool pred);
d4 mfma_f64_16x16x4(double a, double b, d4 c);
f4 mfma_f32_16x16x4(float a, float b, f4 c);

inline unsigned lane_id() { return g_thread.x & 63u; }
// Set bits of `mask` in the lanes below this one (of lanes 0..31 for the
// low word, 32..63 for the high word), plus `acc`.
inline unsigned mbcnt_lo(unsigned mask, unsigned acc) {
    const unsigned l = lane_id();
    const unsigned below = l >= 32 ? mask : (mask & ((1u << l) - 1u));
    return acc + (unsigned)__builtin_popcount(below);
}
inline unsigned mbcnt_hi(unsigned mask, unsigned acc) {
    const unsigned l = lane_id();
    const unsigned below = l < 32 ? 0u : (mask & ((1u << (l - 32)) - 1u));
    return acc + (unsigned)__builtin_popcount(below);
}

}  // namespace emu

#define threadIdx (emu::g_thread)
#define blockIdx (emu::g_block)
#define blockDim (emu::g_block_dim)
#define gridDim (emu::g_grid_dim)

#define __syncthreads() emu::syncthreads()
#define __builtin_amdgcn_readlane(v, l) emu::readlane((v), (l))
#define __builtin_amdgcn_readfirstlane(v) emu::readfirstlane((v))
#define __builtin_amdgcn_update_dpp(o, s, c, rm, bm, bc) emu::update_dpp((o), (s), (c), (rm), (bm), (bc))
#define __builtin_amdgcn_mbcnt_lo(m, a) emu::mbcnt_lo((m), (a))
#define __builtin_amdgcn_mbcnt_hi(m, a) emu::mbcnt_hi((m), (a))
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) emu::mfma_f64_16x16x4((a), (b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) emu::mfma_f32_16x16x4((a), (b), (c))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)

inline unsigned long long __ballot(int pred) { return emu::ballot(pred != 0); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline unsigned __float_as_uint(float x) { unsigned u; std::memcpy(&u, &x, 4); return u; }
inline int __float_as_int(float x) { int u; std::memcpy(&u, &x, 4); return u; }
inline float __int_as_float(int x) { float f; std::memcpy(&f, &x, 4); return f; }
inline long long __double_as_longlong(double x) { long long u; std::memcpy(&u, &x, 8); return u; }
inline double __longlong_as_double(long long x) { double f; std::memcpy(&f, &x, 8); return f; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
// Lanes are fibers of one host thread and never interleave inside a call.
inline unsigned atomicAdd(unsigned *p, unsigned v) {
    const unsigned old = *p;
    *p = old + v;
    return old;
}
