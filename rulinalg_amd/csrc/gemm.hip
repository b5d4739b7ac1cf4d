// GEMM on the gfx950 matrix cores.
//
// One kernel family serves `&a * &b`, the LU Schur update, the Cholesky
// SYRK update and the mixed-precision updates:
//   C = alpha * A * op(B) + beta * C, all row-major with row strides.
//
// Tiling (FP64/FP32): a 256-thread workgroup owns a 128 x 128 tile of C and
// walks K in steps of 16 (FP64) or 32 (FP32), double-buffered through LDS.  The four wave64s
// are arranged 2 x 2; each owns a 64 x 64 sub-tile held as 4 x 4 MFMA
// accumulators of 16 x 16 (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
// A is staged k-major (As[k][i]) so that an MFMA A-fragment (lane l holds
// A[l%16][l/16]) is four contiguous 128-byte row segments; B is staged
// As-is (Bs[k][j]).  The +1 padding of transposed stages spreads the
// transposing LDS writes over all banks.
//
// Low precision (FP16/BF16): operands stay FP32 in memory and are rounded on
// their way into LDS; v_mfma_f32_16x16x32_{f16,bf16} accumulate in FP32.
// Fragments hold 8 consecutive k per lane, so both stages are k-contiguous
// ([i][k] and [j][k]).
#include "kernels.hpp"
#include "mfma.hpp"

namespace rlg {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

constexpr int GBM = 128, GBN = 128, GBK = 16, GTHREADS = 256;

// Maps the linear workgroup id onto (tile_m, tile_n).  Consecutive ids are
// dispatched round-robin over the 8 XCDs; regrouping gives each XCD a
// contiguous range of tiles, and within a range tiles go down columns of
// GROUP_M tile-rows, so that an XCD's L2 sees reuse of both A and B panels.
// The lower-triangle form does not regroup: its tiles above the diagonal
// return at once, so contiguous ranges would hand the XCDs unequal work.
template <bool LOWER>
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int &tm, int &tn) {
    const int total = tiles_m * tiles_n;
    int id = blockIdx.x;
    // the first 8 * floor(total / 8) ids are regrouped; the few left over
    // keep their own tile
    const int full = total & ~7;
    if (!LOWER && total >= 64 && id < full) {
        const int per_xcd = full / 8;
        id = (id % 8) * per_xcd + id / 8;
    }
    constexpr int GROUP_M = 8;
    const int group_size = GROUP_M * tiles_n;
    const int g = id / group_size;
    const int first_m = g * GROUP_M;
    const int gm = min(tiles_m - first_m, GROUP_M);
    const int r = id - g * group_size;
    tm = first_m + r % gm;
    tn = r / gm;
}

// k per LDS stage: 16 in FP64 (2 x 33 KB of stages, two workgroups per CU),
// 32 in FP32 (the same bytes): twice the matrix-core work between barriers.
template <class T> struct StageK { static constexpr int value = GBK; };
template <> struct StageK<float> { static constexpr int value = 2 * GBK; };

// VEC: operands 16-byte aligned with row strides, K (and N for the
// untransposed B) multiples of 16 bytes' worth of elements: the tiles are
// loaded with 16-byte loads (a quarter (FP32) or half (FP64) of the load
// instructions), which measured 18 % faster on the LU's depth-256 FP32
// update.  The untransposed B stage is padded by 16 bytes per row.
// CP: the C tile is read during the last stage (FP32 with VEC only; chosen
// for shallow products, where the epilogue is a large share of a tile's time).
template <class T, bool TRANSB, bool LOWER, bool VEC, bool CP>
__global__ __launch_bounds__(GTHREADS, 2) void k_gemm(
    int M, int N, int K, T alpha, const T *__restrict__ A, long lda,
    const T *__restrict__ B, long ldb, T beta, T *__restrict__ C, long ldc,
    int tiles_m, int tiles_n) {
    using Acc = typename Mfma<T>::acc;
    constexpr int BK = StageK<T>::value;
    constexpr int PER = BK * GBM / GTHREADS;  // staged elements per thread per operand
    constexpr int VL = 16 / sizeof(T);        // elements per 16-byte load
    constexpr int VPER = PER / VL;            // 16-byte loads per thread per operand
    static_assert(GBM == GBN && PER * GTHREADS == BK * GBM && VPER * VL == PER, "whole stages");
    typedef T vT __attribute__((ext_vector_type(VL)));
    __shared__ T As[2][BK][GBM + 1];
    __shared__ __attribute__((aligned(16))) T Bs[2][BK][GBN + (TRANSB ? 1 : VL)];

    int tm, tn;
    tile_coords<LOWER>(tiles_m, tiles_n, tm, tn);
    if (LOWER && tn > tm) return;  // strictly above the diagonal
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;

    Acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    T ra[VEC ? 1 : PER], rb[VEC ? 1 : PER];
    vT rav[VEC ? VPER : 1], rbv[VEC ? VPER : 1];
    auto load = [&](int k0) {
        if constexpr (VEC) {
#pragma unroll
            for (int q = 0; q < VPER; ++q) {
                const int e = q * GTHREADS + tid;
                {   // A tile, 128 x BK in BK / VL vectors per row
                    const int i = e / (BK / VL), gk = k0 + (e % (BK / VL)) * VL, gi = m0 + i;
                    rav[q] = (gi < M && gk < K) ? *reinterpret_cast<const vT *>(A + (long)gi * lda + gk)
                                                : vT{};
                }
                if (!TRANSB) {  // B tile, BK x 128 in 128 / VL vectors per row
                    const int gk = k0 + e / (GBN / VL), gj = n0 + (e % (GBN / VL)) * VL;
                    rbv[q] = (gk < K && gj < N) ? *reinterpret_cast<const vT *>(B + (long)gk * ldb + gj)
                                                : vT{};
                } else {        // B^T given as N x K, vectors along k
                    const int gj = n0 + e / (BK / VL), gk = k0 + (e % (BK / VL)) * VL;
                    rbv[q] = (gj < N && gk < K) ? *reinterpret_cast<const vT *>(B + (long)gj * ldb + gk)
                                                : vT{};
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = q * GTHREADS + tid;
                {   // A tile, 128 x BK: e -> (i = e / BK, kk = e % BK)
                    const int i = e / BK, kk = e % BK;
                    const int gi = m0 + i, gk = k0 + kk;
                    ra[q] = (gi < M && gk < K) ? A[(long)gi * lda + gk] : T(0);
                }
                if (!TRANSB) {  // B tile, BK x 128: e -> (kk = e / 128, j = e % 128)
                    const int kk = e / GBN, j = e % GBN;
                    const int gk = k0 + kk, gj = n0 + j;
                    rb[q] = (gk < K && gj < N) ? B[(long)gk * ldb + gj] : T(0);
                } else {        // B^T given as N x K: e -> (j = e / BK, kk = e % BK)
                    const int j = e / BK, kk = e % BK;
                    const int gj = n0 + j, gk = k0 + kk;
                    rb[q] = (gj < N && gk < K) ? B[(long)gj * ldb + gk] : T(0);
                }
            }
        }
    };
    auto store = [&](int buf) {
        if constexpr (VEC) {
#pragma unroll
            for (int q = 0; q < VPER; ++q) {
                const int e = q * GTHREADS + tid;
                const int i = e / (BK / VL), kb = (e % (BK / VL)) * VL;
#pragma unroll
                for (int v = 0; v < VL; ++v) As[buf][kb + v][i] = rav[q][v];
                if (!TRANSB) {
                    *reinterpret_cast<vT *>(&Bs[buf][e / (GBN / VL)][(e % (GBN / VL)) * VL]) = rbv[q];
                } else {
#pragma unroll
                    for (int v = 0; v < VL; ++v) Bs[buf][kb + v][i] = rbv[q][v];
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = q * GTHREADS + tid;
                As[buf][e % BK][e / BK] = ra[q];
                if (!TRANSB)
                    Bs[buf][e / GBN][e % GBN] = rb[q];
                else
                    Bs[buf][e % BK][e / BK] = rb[q];
            }
        }
    };

    // FP32 with 16-byte tile loads: the first two row groups of the C tile
    // are read during the last stage's MFMAs (32 more registers, within the
    // two-workgroups-per-CU budget; all four, or the scalar-load form, would
    // spill), the other two before the first store.  Measured: depth-256
    // updates 83 -> 89 TF/s; at depth 8192 it costs 6 %, hence CP.
    constexpr bool CPRE = sizeof(T) == 4 && VEC && CP;
    constexpr int CG = CPRE ? 2 : 1;
    T cpre[CG][4][4];
    auto load_c = [&]() {
#pragma unroll
        for (int i = 0; i < CG; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gi = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
                    const int gj = n0 + wn + j * 16 + lr;
                    cpre[i][j][r] = (beta != T(0) && gi < M && gj < N) ? C[(long)gi * ldc + gj] : T(0);
                }
    };

    const int KT = (K + BK - 1) / BK;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) load((kt + 1) * BK);
        else if constexpr (CPRE) load_c();
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            T af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = As[buf][ks * 4 + lk][wm + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][ks * 4 + lk][wn + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
        }
        if (kt + 1 < KT) store(buf ^ 1);
        __syncthreads();
    }
    if constexpr (CPRE) {
        if (KT == 0) load_c();  // K = 0: no stage read it
    }

    // epilogue in four row groups: the 16 loads of C of a group are issued
    // before its stores (a load behind a store to C would wait for it); in
    // FP32 groups 0-1 are in cpre and groups 2-3 are loaded before any store
    T clate[CPRE ? 2 : 1][4][4];
    if constexpr (CPRE) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gi = m0 + wm + (i + 2) * 16 + Mfma<T>::row(lk, r);
                    const int gj = n0 + wn + j * 16 + lr;
                    clate[i][j][r] = (beta != T(0) && gi < M && gj < N) ? C[(long)gi * ldc + gj] : T(0);
                }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T cin[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
                const int gj = n0 + wn + j * 16 + lr;
                if constexpr (CPRE) cin[j][r] = i < 2 ? cpre[i < 2 ? i : 0][j][r] : clate[i >= 2 ? i - 2 : 0][j][r];
                else cin[j][r] = (beta != T(0) && gi < M && gj < N) ? C[(long)gi * ldc + gj] : T(0);
            }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
                const int gj = n0 + wn + j * 16 + lr;
                if (gi < M && gj < N) {
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * cin[j][r];
                    C[(long)gi * ldc + gj] = v;
                }
            }
    }
}

// Depth K <= 64 (the LU and Cholesky updates inside an outer block, which
// are many small launches): the whole of op(B) and A fits one LDS stage, so
// every load of A, B and C is issued before the first wait and a workgroup
// pays one memory latency, where k_gemm pays one per stage plus the
// epilogue's.  Same fragments, k order and epilogue as k_gemm, so the
// results are identical.  Needs 16-byte vectors (the VEC conditions).
constexpr int SHK = 64;
template <class T, bool TRANSB>
__global__ __launch_bounds__(GTHREADS) void k_gemm_shallow(int M, int N, int K, T alpha,
                                                           const T *__restrict__ A, long lda,
                                                           const T *__restrict__ B, long ldb, T beta,
                                                           T *__restrict__ C, long ldc, int tiles_m,
                                                           int tiles_n) {
    using Acc = typename Mfma<T>::acc;
    constexpr int VL = 16 / sizeof(T);
    constexpr int VPER = SHK * GBM / GTHREADS / VL;  // 16-byte loads per thread per operand
    typedef T vT __attribute__((ext_vector_type(VL)));
    __shared__ T As[SHK][GBM + 1];
    __shared__ __attribute__((aligned(16))) T Bs[SHK][GBN + (TRANSB ? 1 : VL)];

    int tm, tn;
    tile_coords<false>(tiles_m, tiles_n, tm, tn);
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;

    vT rav[VPER], rbv[VPER];
#pragma unroll
    for (int q = 0; q < VPER; ++q) {
        const int e = q * GTHREADS + tid;
        {   // A tile, 128 x SHK in SHK / VL vectors per row
            const int i = e / (SHK / VL), gk = (e % (SHK / VL)) * VL, gi = m0 + i;
            rav[q] = (gi < M && gk < K) ? *reinterpret_cast<const vT *>(A + (long)gi * lda + gk) : vT{};
        }
        if (!TRANSB) {  // B tile, SHK x 128 in 128 / VL vectors per row
            const int gk = e / (GBN / VL), gj = n0 + (e % (GBN / VL)) * VL;
            rbv[q] = (gk < K && gj < N) ? *reinterpret_cast<const vT *>(B + (long)gk * ldb + gj) : vT{};
        } else {        // B^T given as N x K, vectors along k
            const int gj = n0 + e / (SHK / VL), gk = (e % (SHK / VL)) * VL;
            rbv[q] = (gj < N && gk < K) ? *reinterpret_cast<const vT *>(B + (long)gj * ldb + gk) : vT{};
        }
    }
    T cin[4][4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
                const int gj = n0 + wn + j * 16 + lr;
                cin[i][j][r] = (beta != T(0) && gi < M && gj < N) ? C[(long)gi * ldc + gj] : T(0);
            }
#pragma unroll
    for (int q = 0; q < VPER; ++q) {
        const int e = q * GTHREADS + tid;
        const int i = e / (SHK / VL), kb = (e % (SHK / VL)) * VL;
#pragma unroll
        for (int v = 0; v < VL; ++v) As[kb + v][i] = rav[q][v];
        if (!TRANSB) {
            *reinterpret_cast<vT *>(&Bs[e / (GBN / VL)][(e % (GBN / VL)) * VL]) = rbv[q];
        } else {
#pragma unroll
            for (int v = 0; v < VL; ++v) Bs[kb + v][i] = rbv[q][v];
        }
    }
    __syncthreads();

    Acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Acc{0, 0, 0, 0};
    // as many k-steps as k_gemm's whole stages take (rows from K on are
    // zero), so that even the signs of zero results agree with it
    constexpr int BK = StageK<T>::value;
    const int ksteps = (K + BK - 1) / BK * (BK / 4);
    for (int ks = 0; ks < ksteps; ++ks) {
        T af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) af[i] = As[ks * 4 + lk][wm + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < 4; ++j) bf[j] = Bs[ks * 4 + lk][wn + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(af[i], bf[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm + i * 16 + Mfma<T>::row(lk, r);
                const int gj = n0 + wn + j * 16 + lr;
                if (gi < M && gj < N) {
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * cin[i][j][r];
                    C[(long)gi * ldc + gj] = v;
                }
            }
}

// ---------------------------------------------------------------- low precision

template <class H> struct LowpMfma;
template <> struct LowpMfma<_Float16> {
    using vec = h8;
    static __device__ __forceinline__ f4 op(vec a, vec b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ _Float16 cvt(float x) { return (_Float16)x; }
};
template <> struct LowpMfma<__bf16> {
    using vec = b8;
    static __device__ __forceinline__ f4 op(vec a, vec b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ __bf16 cvt(float x) { return (__bf16)x; }
};

// C = alpha * acc + beta * C for one wave's 64 x 64 sub-tile at (row0, col0),
// FP32 accumulator layout.  Four row groups; the 16 loads of C of a group are
// issued before its stores (a load behind a store to C would wait for it).
__device__ __forceinline__ void epilogue_f32(const f4 (&acc)[4][4], float *__restrict__ C, long ldc,
                                             int M, int N, int row0, int col0, int lr, int lk,
                                             float alpha, float beta) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float cin[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = row0 + i * 16 + lk * 4 + r, gj = col0 + j * 16 + lr;
                cin[j][r] = (beta != 0.f && gi < M && gj < N) ? C[(long)gi * ldc + gj] : 0.f;
            }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = row0 + i * 16 + lk * 4 + r, gj = col0 + j * 16 + lr;
                if (gi < M && gj < N) {
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.f) v += beta * cin[j][r];
                    C[(long)gi * ldc + gj] = v;
                }
            }
    }
}

constexpr int LBK = 64;          // k per stage (two MFMA k-steps of 32)
constexpr int LPAD = 8;          // halves of padding per staged row

template <class H, bool TRANSB, bool LOWER>
__global__ __launch_bounds__(GTHREADS) void k_gemm_lowp(
    int M, int N, int K, float alpha, const float *__restrict__ A, long lda,
    const float *__restrict__ B, long ldb, float beta, float *__restrict__ C,
    long ldc, int tiles_m, int tiles_n) {
    using V = typename LowpMfma<H>::vec;
    __shared__ __attribute__((aligned(16))) H As[2][GBM][LBK + LPAD];
    __shared__ __attribute__((aligned(16))) H Bs[2][GBN][LBK + LPAD];

    int tm, tn;
    tile_coords<LOWER>(tiles_m, tiles_n, tm, tn);
    if (LOWER && tn > tm) return;
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};

    // 128 x 64 elements per operand per stage = 32 per thread.
    float ra[32], rb[32];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int e = q * GTHREADS + tid;
            {   // A: e -> (i = e / 64, kk = e % 64)
                const int i = e >> 6, kk = e & 63;
                const int gi = m0 + i, gk = k0 + kk;
                ra[q] = (gi < M && gk < K) ? A[(long)gi * lda + gk] : 0.f;
            }
            if (!TRANSB) {  // B: e -> (kk = e / 128, j = e % 128)
                const int kk = e >> 7, j = e & 127;
                const int gk = k0 + kk, gj = n0 + j;
                rb[q] = (gk < K && gj < N) ? B[(long)gk * ldb + gj] : 0.f;
            } else {
                const int j = e >> 6, kk = e & 63;
                const int gj = n0 + j, gk = k0 + kk;
                rb[q] = (gj < N && gk < K) ? B[(long)gj * ldb + gk] : 0.f;
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int e = q * GTHREADS + tid;
            As[buf][e >> 6][e & 63] = LowpMfma<H>::cvt(ra[q]);
            if (!TRANSB)
                Bs[buf][e & 127][e >> 7] = LowpMfma<H>::cvt(rb[q]);
            else
                Bs[buf][e >> 6][e & 63] = LowpMfma<H>::cvt(rb[q]);
        }
    };

    const int KT = (K + LBK - 1) / LBK;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) load((kt + 1) * LBK);
#pragma unroll
        for (int ks = 0; ks < LBK / 32; ++ks) {
            V af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const V *>(&As[buf][wm + i * 16 + lr][ks * 32 + lk * 8]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bf[j] = *reinterpret_cast<const V *>(&Bs[buf][wn + j * 16 + lr][ks * 32 + lk * 8]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = LowpMfma<H>::op(af[i], bf[j], acc[i][j]);
        }
        if (kt + 1 < KT) store(buf ^ 1);
        __syncthreads();
    }

    epilogue_f32(acc, C, ldc, M, N, m0 + wm, n0 + wn, lr, lk, alpha, beta);
}

// ---------------------------------------------------------------- packed low precision
//
// The packed form rounds each operand to FP16/BF16 once, into k-contiguous
// rows (A as m x Kp, B as n x Kp, with Kp = K rounded up to 64 and the pad
// zero), so the GEMM itself streams 16-byte chunks straight into LDS and
// needs no conversion or transposition in its inner loop.

constexpr int PKB = 64;  // k granularity of the packed operands

// dst[i][k] = src[i][k] (k < K), 0 for K <= k < Kp.  One thread per 8 k.
template <class H>
__global__ __launch_bounds__(256) void k_pack_rows(const float *__restrict__ src, long lds, int rows,
                                                   int K, int Kp, H *__restrict__ dst, long ldd) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int chunks = Kp / 8;
    const long i = e / chunks;
    const int k0 = (int)(e - i * chunks) * 8;
    if (i >= rows) return;
    const float *s = src + i * lds + k0;
    typedef H h8v __attribute__((ext_vector_type(8)));
    h8v v;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = LowpMfma<H>::cvt(k0 + q < K ? s[q] : 0.f);
    *reinterpret_cast<h8v *>(dst + i * ldd + k0) = v;
}

// dst[j][k] = src[k][j] for a K x cols source (k < K), 0 for K <= k < Kp.
// 64 x 64 tiles through LDS: coalesced reads along j, writes along k.
template <class H>
__global__ __launch_bounds__(256) void k_pack_cols(const float *__restrict__ src, long lds, int K,
                                                   int cols, int Kp, H *__restrict__ dst, long ldd) {
    __shared__ float t[64][65];
    const int k0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int kk = e >> 6, jj = e & 63;
        const int k = k0 + kk, j = j0 + jj;
        t[kk][jj] = (k < K && j < cols) ? src[(long)k * lds + j] : 0.f;
    }
    __syncthreads();
    typedef H h8v __attribute__((ext_vector_type(8)));
    for (int e = threadIdx.x; e < 64 * 8; e += 256) {
        const int jj = e >> 3, kc = (e & 7) * 8;
        const int j = j0 + jj;
        if (j >= cols || k0 + kc >= Kp) continue;
        h8v v;
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = LowpMfma<H>::cvt(t[kc + q][jj]);
        *reinterpret_cast<h8v *>(dst + (long)j * ldd + k0 + kc) = v;
    }
}

// C = alpha * Ah * Bh^T + beta * C with Ah (M x Kp) and Bh (N x Kp) packed,
// Kp a multiple of 64.  128 x 128 tile per workgroup, 2 x 2 waves of 64 x 64,
// k staged 64 at a time through double-buffered LDS with 16-byte moves.
template <class H, bool LOWER>
__global__ __launch_bounds__(GTHREADS) void k_gemm_nt_lowp(
    int M, int N, int Kp, float alpha, const H *__restrict__ A, long lda, const H *__restrict__ B,
    long ldb, float beta, float *__restrict__ C, long ldc, int tiles_m, int tiles_n) {
    using V = typename LowpMfma<H>::vec;
    __shared__ __attribute__((aligned(16))) H As[2][GBM][LBK + LPAD];
    __shared__ __attribute__((aligned(16))) H Bs[2][GBN][LBK + LPAD];

    int tm, tn;
    tile_coords<LOWER>(tiles_m, tiles_n, tm, tn);
    if (LOWER && tn > tm) return;
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};

    // a stage is 128 rows x 64 k per operand = 1024 chunks of 8 halves: 4 per thread
    uint4 ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = q * GTHREADS + tid;
            const int row = c >> 3, kc = (c & 7) * 8;
            const int gi = m0 + row, gj = n0 + row;
            ra[q] = gi < M ? *reinterpret_cast<const uint4 *>(A + (long)gi * lda + k0 + kc)
                           : uint4{0, 0, 0, 0};
            rb[q] = gj < N ? *reinterpret_cast<const uint4 *>(B + (long)gj * ldb + k0 + kc)
                           : uint4{0, 0, 0, 0};
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = q * GTHREADS + tid;
            const int row = c >> 3, kc = (c & 7) * 8;
            *reinterpret_cast<uint4 *>(&As[buf][row][kc]) = ra[q];
            *reinterpret_cast<uint4 *>(&Bs[buf][row][kc]) = rb[q];
        }
    };

    const int KT = Kp / LBK;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) load((kt + 1) * LBK);
#pragma unroll
        for (int ks = 0; ks < LBK / 32; ++ks) {
            V af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = *reinterpret_cast<const V *>(&As[buf][wm + i * 16 + lr][ks * 32 + lk * 8]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bf[j] = *reinterpret_cast<const V *>(&Bs[buf][wn + j * 16 + lr][ks * 32 + lk * 8]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = LowpMfma<H>::op(af[i], bf[j], acc[i][j]);
        }
        if (kt + 1 < KT) store(buf ^ 1);
        __syncthreads();
    }

    epilogue_f32(acc, C, ldc, M, N, m0 + wm, n0 + wn, lr, lk, alpha, beta);
}

// ---------------------------------------------------------------- split FP16
//
// FP32 products on the FP16 matrix cores: every FP32 operand x is scaled by
// a power of two per row (so that the row's largest |x| lands in
// [2^13, 2^14)) and split as x = hi + lo, hi = fp16(x), lo = fp16(x - hi);
// the product accumulates hi_a hi_b + hi_a lo_b + lo_a hi_b in FP32 (the
// dropped lo_a lo_b is below 2^-22 of the scaled magnitudes), and the
// epilogue multiplies by the two inverse scales.  The result is as accurate
// as an FP32 product (normwise error of order 2^-22 |A||B| plus FP32
// accumulation) at three FP16 MFMAs per k-step.

constexpr int SBK = 32;  // k per stage of the split kernel (four operands staged)

// Power-of-two scale for a row whose largest |x| is mx, and its inverse.
__device__ __forceinline__ void split_scale(float mx, float &sc, float &inv) {
    if (!(mx > 0.f) || !isfinite(mx)) { sc = inv = 1.f; return; }
    int e;
    frexpf(mx, &e);                      // mx = f 2^e, f in [0.5, 1)
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    sc = ldexpf(1.f, 14 - e);
    inv = ldexpf(1.f, e - 14);
}

__device__ __forceinline__ void split_one(float x, float sc, _Float16 &hi, _Float16 &lo) {
    const float y = x * sc;
    hi = (_Float16)y;
    lo = (_Float16)(y - (float)hi);
}

// dst_{hi,lo}[i][k] from the rows of src (rows x K, k < K; zero up to Kp),
// one wave per row; inv[i] = the row's inverse scale.
__global__ __launch_bounds__(256) void k_pack_rows_split(const float *__restrict__ src, long lds, int rows,
                                                         int K, int Kp, _Float16 *__restrict__ hi,
                                                         _Float16 *__restrict__ lo, long ldd,
                                                         float *__restrict__ inv) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= rows) return;
    const float *r = src + (long)i * lds;
    float mx = 0.f;
    for (int k = lane; k < K; k += 64) mx = fmaxf(mx, fabsf(r[k]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    float sc, iv;
    split_scale(mx, sc, iv);
    if (lane == 0) inv[i] = iv;
    typedef _Float16 h8v __attribute__((ext_vector_type(8)));
    for (int k0 = lane * 8; k0 < Kp; k0 += 64 * 8) {
        h8v vh, vl;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            _Float16 a, b;
            split_one(k0 + q < K ? r[k0 + q] : 0.f, sc, a, b);
            vh[q] = a;
            vl[q] = b;
        }
        *reinterpret_cast<h8v *>(hi + (long)i * ldd + k0) = vh;
        *reinterpret_cast<h8v *>(lo + (long)i * ldd + k0) = vl;
    }
}

// dst_{hi,lo}[j][k] = split(src[k][j]) for a K x cols source, 64 columns per
// workgroup: the column maxima over all K first, then 64 x 64 tiles through
// LDS (coalesced reads along j, writes along k).
__global__ __launch_bounds__(256) void k_pack_cols_split(const float *__restrict__ src, long lds, int K,
                                                         int cols, int Kp, _Float16 *__restrict__ hi,
                                                         _Float16 *__restrict__ lo, long ldd,
                                                         float *__restrict__ inv) {
    __shared__ float t[64][65];
    __shared__ float cm[4][64];
    __shared__ float csc[64];
    const int j0 = blockIdx.x * 64;
    const int jj = threadIdx.x & 63, kq = threadIdx.x >> 6;
    {
        float mx = 0.f;
        const int j = j0 + jj;
        if (j < cols)
            for (int k = kq; k < K; k += 4) mx = fmaxf(mx, fabsf(src[(long)k * lds + j]));
        cm[kq][jj] = mx;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const float mx = fmaxf(fmaxf(cm[0][jj], cm[1][jj]), fmaxf(cm[2][jj], cm[3][jj]));
        float sc, iv;
        split_scale(mx, sc, iv);
        csc[jj] = sc;
        if (j0 + jj < cols) inv[j0 + jj] = iv;
    }
    typedef _Float16 h8v __attribute__((ext_vector_type(8)));
    for (int k0 = 0; k0 < Kp; k0 += 64) {
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int kk = e >> 6, c = e & 63;
            const int k = k0 + kk, j = j0 + c;
            t[kk][c] = (k < K && j < cols) ? src[(long)k * lds + j] : 0.f;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 8; e += 256) {
            const int c = e >> 3, kc = (e & 7) * 8;
            const int j = j0 + c;
            if (j >= cols) continue;
            h8v vh, vl;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                _Float16 a, b;
                split_one(t[kc + q][c], csc[c], a, b);
                vh[q] = a;
                vl[q] = b;
            }
            *reinterpret_cast<h8v *>(hi + (long)j * ldd + k0 + kc) = vh;
            *reinterpret_cast<h8v *>(lo + (long)j * ldd + k0 + kc) = vl;
        }
    }
}

// C = alpha * (Ah Bh^T + Ah Bl^T + Al Bh^T) diag-scaled + beta * C with the
// split operands (M x Kp and N x Kp, Kp a multiple of 64): 128 x 128 tile per
// workgroup, 2 x 2 waves of 64 x 64, four operands staged 32 k at a time
// through double-buffered LDS with 16-byte moves.
template <bool LOWER>
__global__ __launch_bounds__(GTHREADS, 2) void k_gemm_nt_split(
    int M, int N, int Kp, float alpha, const _Float16 *__restrict__ Ah, const _Float16 *__restrict__ Al,
    const float *__restrict__ ia, const _Float16 *__restrict__ Bh, const _Float16 *__restrict__ Bl,
    const float *__restrict__ ib, long ldk, float beta, float *__restrict__ C, long ldc, int tiles_m,
    int tiles_n) {
    using V = h8;
    __shared__ __attribute__((aligned(16))) _Float16 S[2][4][GBM][SBK + LPAD];

    int tm, tn;
    tile_coords<LOWER>(tiles_m, tiles_n, tm, tn);
    if (LOWER && tn > tm) return;
    const int m0 = tm * GBM, n0 = tn * GBN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};

    // a stage is 128 rows x 32 k per operand = 512 chunks of 8 halves: 2 per thread
    uint4 rg[4][2];
    const _Float16 *src[4] = {Ah, Al, Bh, Bl};
    auto load = [&](int k0) {
#pragma unroll
        for (int o = 0; o < 4; ++o)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int c = q * GTHREADS + tid;
                const int row = c >> 2, kc = (c & 3) * 8;
                const int g = (o < 2 ? m0 : n0) + row;
                const bool in = g < (o < 2 ? M : N);
                rg[o][q] = in ? *reinterpret_cast<const uint4 *>(src[o] + (long)g * ldk + k0 + kc)
                              : uint4{0, 0, 0, 0};
            }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int o = 0; o < 4; ++o)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int c = q * GTHREADS + tid;
                const int row = c >> 2, kc = (c & 3) * 8;
                *reinterpret_cast<uint4 *>(&S[buf][o][row][kc]) = rg[o][q];
            }
    };

    const int KT = Kp / SBK;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < KT) load((kt + 1) * SBK);
        V ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ah[i] = *reinterpret_cast<const V *>(&S[buf][0][wm + i * 16 + lr][lk * 8]);
            al[i] = *reinterpret_cast<const V *>(&S[buf][1][wm + i * 16 + lr][lk * 8]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bh[j] = *reinterpret_cast<const V *>(&S[buf][2][wn + j * 16 + lr][lk * 8]);
            bl[j] = *reinterpret_cast<const V *>(&S[buf][3][wn + j * 16 + lr][lk * 8]);
        }
        // the small cross terms first, then the leading one
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        if (kt + 1 < KT) store(buf ^ 1);
        __syncthreads();
    }

    // undo the row and column scales (powers of two: exact)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gj = n0 + wn + j * 16 + lr;
            const float sj = gj < N ? ib[gj] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm + i * 16 + lk * 4 + r;
                const float si = gi < M ? ia[gi] : 0.f;
                acc[i][j][r] = acc[i][j][r] * si * sj;
            }
        }
    epilogue_f32(acc, C, ldc, M, N, m0 + wm, n0 + wn, lr, lk, alpha, beta);
}

static size_t al256z(size_t x) { return (x + 255) & ~(size_t)255; }

size_t gemm_lowp_pack_bytes(size_t m, size_t n, size_t k) {
    // room for the split form: hi and lo copies of both operands and their
    // inverse scales (the plain FP16/BF16 form uses the first half)
    const size_t kp = (k + PKB - 1) / PKB * PKB;
    return 2 * al256z(m * kp * 2) + 2 * al256z(n * kp * 2) + al256z(m * 4) + al256z(n * 4);
}

static int gemm_split_packed(hipStream_t s, size_t m, size_t n, size_t k, float alpha, const float *a,
                             size_t lda, const float *b, size_t ldb, float beta, float *c, size_t ldc,
                             bool trans_b, bool lower_only, void *pack, size_t lds_pad) {
    const size_t kp = (k + PKB - 1) / PKB * PKB;
    char *p = static_cast<char *>(pack);
    _Float16 *ah = reinterpret_cast<_Float16 *>(p); p += al256z(m * kp * 2);
    _Float16 *al = reinterpret_cast<_Float16 *>(p); p += al256z(m * kp * 2);
    _Float16 *bh = reinterpret_cast<_Float16 *>(p); p += al256z(n * kp * 2);
    _Float16 *bl = reinterpret_cast<_Float16 *>(p); p += al256z(n * kp * 2);
    float *ia = reinterpret_cast<float *>(p); p += al256z(m * 4);
    float *ib = reinterpret_cast<float *>(p);
    {
        RLG_PROF_W("pack_split", s, 4.0 * (double)(m + n) * k + 4.0 * (double)(m + n) * kp);
        hipLaunchKernelGGL(k_pack_rows_split, dim3(ceil_div(m, 4)), dim3(256), 0, s, a, (long)lda, (int)m,
                           (int)k, (int)kp, ah, al, (long)kp, ia);
        RLG_LAUNCHED();
        if (trans_b)
            hipLaunchKernelGGL(k_pack_rows_split, dim3(ceil_div(n, 4)), dim3(256), 0, s, b, (long)ldb,
                               (int)n, (int)k, (int)kp, bh, bl, (long)kp, ib);
        else
            hipLaunchKernelGGL(k_pack_cols_split, dim3(ceil_div(n, 64)), dim3(256), 0, s, b, (long)ldb,
                               (int)k, (int)n, (int)kp, bh, bl, (long)kp, ib);
        RLG_LAUNCHED();
    }
    const int tiles_m = (int)ceil_div(m, GBM), tiles_n = (int)ceil_div(n, GBN);
    const double flops = (lower_only ? 1.0 : 2.0) * (double)m * (double)n * (double)k;
    RLG_PROF_W("gemm_f16x3", s, flops);
    if (lower_only)
        hipLaunchKernelGGL(k_gemm_nt_split<true>, dim3(tiles_m * tiles_n), dim3(GTHREADS), (unsigned)lds_pad,
                           s, (int)m, (int)n, (int)kp, alpha, ah, al, ia, bh, bl, ib, (long)kp, beta, c,
                           (long)ldc, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL(k_gemm_nt_split<false>, dim3(tiles_m * tiles_n), dim3(GTHREADS), (unsigned)lds_pad,
                           s, (int)m, (int)n, (int)kp, alpha, ah, al, ia, bh, bl, ib, (long)kp, beta, c,
                           (long)ldc, tiles_m, tiles_n);
    RLG_LAUNCHED();
    return RLG_OK;
}

template <class H>
static int gemm_lowp_packed(hipStream_t s, const char *cls, size_t m, size_t n, size_t k, float alpha,
                            const float *a, size_t lda, const float *b, size_t ldb, float beta,
                            float *c, size_t ldc, bool trans_b, bool lower_only, void *pack) {
    const size_t kp = (k + PKB - 1) / PKB * PKB;
    H *ah = static_cast<H *>(pack);
    H *bh = reinterpret_cast<H *>(static_cast<char *>(pack) + ((m * kp * 2 + 255) & ~(size_t)255));
    {
        RLG_PROF_W("pack_lowp", s, 4.0 * (double)(m + n) * k + 2.0 * (double)(m + n) * kp);
        hipLaunchKernelGGL(k_pack_rows<H>, dim3(ceil_div(m * (kp / 8), 256)), dim3(256), 0, s, a,
                           (long)lda, (int)m, (int)k, (int)kp, ah, (long)kp);
        RLG_LAUNCHED();
        if (trans_b)
            hipLaunchKernelGGL(k_pack_rows<H>, dim3(ceil_div(n * (kp / 8), 256)), dim3(256), 0, s, b,
                               (long)ldb, (int)n, (int)k, (int)kp, bh, (long)kp);
        else
            hipLaunchKernelGGL(k_pack_cols<H>, dim3(ceil_div(n, 64), ceil_div(kp, 64)), dim3(256), 0,
                               s, b, (long)ldb, (int)k, (int)n, (int)kp, bh, (long)kp);
        RLG_LAUNCHED();
    }
    const int tiles_m = (int)ceil_div(m, GBM), tiles_n = (int)ceil_div(n, GBN);
    const double flops = (lower_only ? 1.0 : 2.0) * (double)m * (double)n * (double)k;
    static const int cls_id = prof_class(cls);
    ProfScope prof_scope(cls_id, s, flops);
    if (lower_only)
        hipLaunchKernelGGL((k_gemm_nt_lowp<H, true>), dim3(tiles_m * tiles_n), dim3(GTHREADS), 0, s,
                           (int)m, (int)n, (int)kp, alpha, ah, (long)kp, bh, (long)kp, beta, c,
                           (long)ldc, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL((k_gemm_nt_lowp<H, false>), dim3(tiles_m * tiles_n), dim3(GTHREADS), 0, s,
                           (int)m, (int)n, (int)kp, alpha, ah, (long)kp, bh, (long)kp, beta, c,
                           (long)ldc, tiles_m, tiles_n);
    RLG_LAUNCHED();
    return RLG_OK;
}

// ---------------------------------------------------------------- launchers

static int check_dims(size_t m, size_t n, size_t k) {
    if (m > (size_t)INT32_MAX / 2 || n > (size_t)INT32_MAX / 2 || k > (size_t)INT32_MAX / 2) {
        set_error("gemm dimension too large (%zu, %zu, %zu)", m, n, k);
        return RLG_E_INVALID_ARG;
    }
    return RLG_OK;
}

template <class T>
static int gemm_impl(hipStream_t s, const char *cls, size_t m, size_t n, size_t k, T alpha,
                     const T *a, size_t lda, const T *b, size_t ldb, T beta, T *c,
                     size_t ldc, bool trans_b, bool lower_only, size_t lds_pad = 0) {
    if (m == 0 || n == 0) return RLG_OK;
    RLG_TRY(check_dims(m, n, k));
    const int tiles_m = (int)ceil_div(m, GBM), tiles_n = (int)ceil_div(n, GBN);
    dim3 grid(tiles_m * tiles_n), block(GTHREADS);
    const double flops = (lower_only ? 1.0 : 2.0) * (double)m * (double)n * (double)k;
    static const int cls_id = prof_class(cls);
    ProfScope prof_scope(cls_id, s, flops);
    // 16-byte loads when every vector the kernel reads is whole and aligned
    constexpr size_t VL = 16 / sizeof(T);
    auto al16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const bool vec = k % VL == 0 && lda % VL == 0 && ldb % VL == 0 && al16(a) && al16(b) &&
                     (trans_b || n % VL == 0);
    const bool cp = vec && sizeof(T) == 4 && k <= 1024;
    if (vec && !lower_only && k > 0 && k <= (size_t)SHK && lds_pad == 0) {
        if (trans_b)
            hipLaunchKernelGGL((k_gemm_shallow<T, true>), grid, block, 0, s, (int)m, (int)n, (int)k, alpha,
                               a, (long)lda, b, (long)ldb, beta, c, (long)ldc, tiles_m, tiles_n);
        else
            hipLaunchKernelGGL((k_gemm_shallow<T, false>), grid, block, 0, s, (int)m, (int)n, (int)k, alpha,
                               a, (long)lda, b, (long)ldb, beta, c, (long)ldc, tiles_m, tiles_n);
        RLG_LAUNCHED();
        return RLG_OK;
    }
#define RLG_GEMM_LAUNCH(TB, LO, V, CPF)                                                                \
    hipLaunchKernelGGL((k_gemm<T, TB, LO, V, CPF>), grid, block, (unsigned)lds_pad, s, (int)m, (int)n, \
                       (int)k, alpha, a, (long)lda, b, (long)ldb, beta, c, (long)ldc, tiles_m, tiles_n)
#define RLG_GEMM_LAUNCH_V(TB, LO)                                                   \
    do {                                                                            \
        if (cp) RLG_GEMM_LAUNCH(TB, LO, true, true);                                \
        else if (vec) RLG_GEMM_LAUNCH(TB, LO, true, false);                         \
        else RLG_GEMM_LAUNCH(TB, LO, false, false);                                 \
    } while (0)
    if (trans_b) {
        if (lower_only) RLG_GEMM_LAUNCH_V(true, true); else RLG_GEMM_LAUNCH_V(true, false);
    } else {
        if (lower_only) RLG_GEMM_LAUNCH_V(false, true); else RLG_GEMM_LAUNCH_V(false, false);
    }
#undef RLG_GEMM_LAUNCH_V
#undef RLG_GEMM_LAUNCH
    RLG_LAUNCHED();
    return RLG_OK;
}

int gemm_f64(hipStream_t s, size_t m, size_t n, size_t k, double alpha, const double *a,
             size_t lda, const double *b, size_t ldb, double beta, double *c, size_t ldc,
             bool trans_b, bool lower_only) {
    return gemm_impl<double>(s, "gemm_f64", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                             trans_b, lower_only);
}

int gemm_f32(hipStream_t s, size_t m, size_t n, size_t k, float alpha, const float *a,
             size_t lda, const float *b, size_t ldb, float beta, float *c, size_t ldc,
             bool trans_b, bool lower_only) {
    return gemm_impl<float>(s, "gemm_f32", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                            trans_b, lower_only);
}

int gemm_f64_padded(hipStream_t s, size_t m, size_t n, size_t k, double alpha, const double *a,
                    size_t lda, const double *b, size_t ldb, double beta, double *c, size_t ldc,
                    bool trans_b, bool lower_only, size_t lds_pad) {
    return gemm_impl<double>(s, "gemm_f64", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, trans_b,
                             lower_only, lds_pad);
}

int gemm_f32_padded(hipStream_t s, size_t m, size_t n, size_t k, float alpha, const float *a,
                    size_t lda, const float *b, size_t ldb, float beta, float *c, size_t ldc,
                    bool trans_b, bool lower_only, size_t lds_pad) {
    return gemm_impl<float>(s, "gemm_f32", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, trans_b,
                            lower_only, lds_pad);
}

int gemm_lowp(hipStream_t s, int prec, size_t m, size_t n, size_t k, float alpha,
              const float *a, size_t lda, const float *b, size_t ldb, float beta, float *c,
              size_t ldc, bool trans_b, bool lower_only, void *pack, size_t pack_bytes,
              size_t lds_pad) {
    if (m == 0 || n == 0) return RLG_OK;
    RLG_TRY(check_dims(m, n, k));
    if (prec == RLG_PREC_F16X3) {
        if (k == 0 || !pack || pack_bytes < gemm_lowp_pack_bytes(m, n, k)) {
            set_error("gemm_lowp: the split FP16 form needs its packed-operand area");
            return RLG_E_INVALID_ARG;
        }
        return gemm_split_packed(s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, trans_b, lower_only, pack,
                                 lds_pad);
    }
    if (k > 0 && pack && pack_bytes >= gemm_lowp_pack_bytes(m, n, k)) {
        if (prec == RLG_PREC_F16)
            return gemm_lowp_packed<_Float16>(s, "gemm_f16", m, n, k, alpha, a, lda, b, ldb, beta, c,
                                              ldc, trans_b, lower_only, pack);
        if (prec == RLG_PREC_BF16)
            return gemm_lowp_packed<__bf16>(s, "gemm_bf16", m, n, k, alpha, a, lda, b, ldb, beta, c,
                                            ldc, trans_b, lower_only, pack);
    }
    const int tiles_m = (int)ceil_div(m, GBM), tiles_n = (int)ceil_div(n, GBN);
    dim3 grid(tiles_m * tiles_n), block(GTHREADS);
#define RLG_LOWP_LAUNCH(H, TB, LO)                                                           \
    hipLaunchKernelGGL((k_gemm_lowp<H, TB, LO>), grid, block, 0, s, (int)m, (int)n, (int)k,  \
                       alpha, a, (long)lda, b, (long)ldb, beta, c, (long)ldc, tiles_m,       \
                       tiles_n)
    const double flops = (lower_only ? 1.0 : 2.0) * (double)m * (double)n * (double)k;
    if (prec == RLG_PREC_F16) {
        RLG_PROF_W("gemm_f16", s, flops);
        if (trans_b) {
            if (lower_only) RLG_LOWP_LAUNCH(_Float16, true, true); else RLG_LOWP_LAUNCH(_Float16, true, false);
        } else {
            if (lower_only) RLG_LOWP_LAUNCH(_Float16, false, true); else RLG_LOWP_LAUNCH(_Float16, false, false);
        }
    } else if (prec == RLG_PREC_BF16) {
        RLG_PROF_W("gemm_bf16", s, flops);
        if (trans_b) {
            if (lower_only) RLG_LOWP_LAUNCH(__bf16, true, true); else RLG_LOWP_LAUNCH(__bf16, true, false);
        } else {
            if (lower_only) RLG_LOWP_LAUNCH(__bf16, false, true); else RLG_LOWP_LAUNCH(__bf16, false, false);
        }
    } else {
        set_error("gemm_lowp: precision %d is not FP16/BF16", prec);
        return RLG_E_INVALID_ARG;
    }
#undef RLG_LOWP_LAUNCH
    RLG_LAUNCHED();
    return RLG_OK;
}

// ---------------------------------------------------------------- GEMV

// y = A x, one wave per row, FP64.
__global__ __launch_bounds__(256) void k_gemv_f64(int m, int n, const double *__restrict__ a,
                                                  long lda, const double *__restrict__ x,
                                                  double *__restrict__ y) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= m) return;
    const double *r = a + (long)row * lda;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += r[j] * x[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[row] = s;
}

int gemv_f64(hipStream_t s, size_t m, size_t n, const double *a, size_t lda, const double *x,
             double *y) {
    if (m == 0) return RLG_OK;
    RLG_TRY(check_dims(m, n, 1));
    RLG_PROF("gemv_f64", s);
    hipLaunchKernelGGL(k_gemv_f64, dim3(ceil_div(m, 4)), dim3(256), 0, s, (int)m, (int)n, a,
                       (long)lda, x, y);
    RLG_LAUNCHED();
    return RLG_OK;
}

}  // namespace rlg
