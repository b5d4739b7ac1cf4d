// 16 x 16 x 4 matrix-core products in FP64 and FP32 on gfx950.
//
// Operand layout (both instructions): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] of one k-step of 4.  Accumulator layout of the 16 x 16
// result: for the FP32 instruction lane l, register r holds
// D[4 (l / 16) + r][l % 16]; for the FP64 instruction it holds
// D[(l / 16) + 4 r][l % 16].
#pragma once

#include <hip/hip_runtime.h>

namespace rlg {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <class T> struct Mfma;
template <> struct Mfma<double> {
    using acc = d4;
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    using acc = f4;
    static __device__ __forceinline__ int row(int lk, int r) { return lk * 4 + r; }
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

}  // namespace rlg
