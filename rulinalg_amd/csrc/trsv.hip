// Triangular solves on nrhs right-hand sides (vector r at x + r * ldx),
// blocked by 64 rows: a diagonal block is solved by one workgroup per RHS
// in LDS (column sweep), and its contribution to the rest of the vector is
// removed by an MFMA GEMM on the transposed RHS block.
//
// The column sweeps follow the reference's substitution orders where the
// reference itself sweeps by columns: transpose_back_substitution
// (src/matrix/decomposition/cholesky.rs:329-361) divides x[i] by l_ii and
// then subtracts x[i] * l[i][j] from x[j] for j < i, exactly as the
// lower-transposed block kernel does.  forward_substitution,
// lu_forward_substitution and back_substitution (src/matrix/mod.rs:318-398,
// lu.rs:624-642) use row dot products in the reference; here the same
// quantities are accumulated column by column.
#include "kernels.hpp"

namespace rlg {

constexpr int TB = 64;

enum TriMode { LOWER_UNIT = 0, LOWER = 1, UPPER = 2, LOWER_T = 3 };

template <class T, int MODE>
__global__ __launch_bounds__(64) void k_trsv_diag(const T *__restrict__ M, long lda, int b0, int bs,
                                                  T *__restrict__ X, long ldx) {
#pragma clang fp contract(off)
    __shared__ T m[TB][TB + 1];
    __shared__ T x[TB];
    const int t = threadIdx.x;
    T *xr = X + (long)blockIdx.x * ldx + b0;
    for (int r = 0; r < bs; ++r) {
        const bool need = (MODE == UPPER) ? (t >= r) : (t <= r);
        if (t < bs && need) m[r][t] = M[(long)(b0 + r) * lda + b0 + t];
    }
    if (t < bs) x[t] = xr[t];
    __syncthreads();
    if (MODE == LOWER_UNIT || MODE == LOWER) {
        for (int c = 0; c < bs; ++c) {
            if (MODE == LOWER && t == c) x[c] = x[c] / m[c][c];
            __syncthreads();
            const T xc = x[c];
            if (t > c && t < bs) x[t] = x[t] - m[t][c] * xc;
            __syncthreads();
        }
    } else if (MODE == UPPER) {
        for (int c = bs - 1; c >= 0; --c) {
            if (t == c) x[c] = x[c] / m[c][c];
            __syncthreads();
            const T xc = x[c];
            if (t < c) x[t] = x[t] - m[t][c] * xc;
            __syncthreads();
        }
    } else {  // LOWER_T: solve L^T y = x with L lower
        for (int c = bs - 1; c >= 0; --c) {
            if (t == c) x[c] = x[c] / m[c][c];
            __syncthreads();
            const T xc = x[c];
            if (t < c) x[t] = x[t] - xc * m[c][t];
            __syncthreads();
        }
    }
    if (t < bs) xr[t] = x[t];
}

template <class T, int MODE>
static int diag(hipStream_t s, const T *m, size_t lda, size_t b0, size_t bs, size_t nrhs, T *x,
                size_t ldx) {
    hipLaunchKernelGGL((k_trsv_diag<T, MODE>), dim3((unsigned)nrhs), dim3(64), 0, s, m, (long)lda,
                       (int)b0, (int)bs, x, (long)ldx);
    RLG_LAUNCHED();
    return RLG_OK;
}

// ---------------------------------------------------------------- one launch
//
// Single-launch blocked substitution.  Workgroups take 64-row blocks in
// solve order from a ticket counter.  The workgroup with ticket t gathers the
// contributions of the blocks solved before it (tickets 0..t-1) tile by tile:
// it loads the 64 x 64 tile of the matrix first, then waits for that block's
// published flag, then accumulates.  Finally one wave per right-hand side
// solves the diagonal block with lane-to-lane broadcasts (no workgroup
// barriers), the block of X is written and its flag released.
//
// A workgroup waits only on blocks whose tickets are smaller than its own;
// those were taken by workgroups that are already running and never wait on
// larger tickets, so every wait ends.  The waits are still bounded: a wait
// that exceeds 2^22 polls raises the timeout flag and the workgroup moves on.
constexpr int FB = 64;
constexpr int FMAXR = 16;

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_trsv_flag(const T *__restrict__ M, long lda, int n, int nrhs,
                                                   T *__restrict__ X, long ldx,
                                                   unsigned *__restrict__ ticket,
                                                   unsigned *__restrict__ flags, int *timeout) {
    __shared__ T tile[FB][FB + 1];
    __shared__ T xs[FMAXR][FB];
    __shared__ unsigned s_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nb = (n + FB - 1) / FB;
    constexpr bool fwd = MODE == LOWER_UNIT || MODE == LOWER;
    for (;;) {
        if (tid == 0) s_t = atomicAdd(ticket, 1u);
        __syncthreads();
        const int t = (int)s_t;
        __syncthreads();
        if (t >= nb) break;
        const int blk = fwd ? t : nb - 1 - t;
        const int b0 = blk * FB, bs = min(FB, n - b0);
        T acc[FMAXR / 4];
#pragma unroll
        for (int q = 0; q < FMAXR / 4; ++q) acc[q] = T(0);
        for (int q = 0; q < t; ++q) {
            const int cb = fwd ? q : nb - 1 - q;
            const int c0 = cb * FB, cs = min(FB, n - c0);
            // tile[i][j] = op(M)[b0 + i][c0 + j]
            for (int e = tid; e < FB * FB; e += 256) {
                const int hi = e >> 6, lo = e & 63;
                if (MODE == LOWER_T) {  // L^T[b0+i][c0+j] = L[c0+j][b0+i]: read rows of L
                    const int j = hi, i = lo;
                    tile[i][j] = (i < bs && j < cs) ? M[(long)(c0 + j) * lda + b0 + i] : T(0);
                } else {
                    const int i = hi, j = lo;
                    tile[i][j] = (i < bs && j < cs) ? M[(long)(b0 + i) * lda + c0 + j] : T(0);
                }
            }
            if (tid == 0) {
                unsigned spins = 0;
                while (__hip_atomic_load(&flags[cb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                    if (++spins > (1u << 22) ||
                        __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            for (int e = tid; e < nrhs * FB; e += 256) {
                const int r = e >> 6, j = e & 63;
                xs[r][j] = j < cs ? X[(long)r * ldx + c0 + j] : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < FMAXR / 4; ++q) {
                const int r = wave + 4 * q;
                if (r < nrhs) {
                    T sacc = acc[q];
                    for (int j = 0; j < FB; ++j) sacc += tile[lane][j] * xs[r][j];
                    acc[q] = sacc;
                }
            }
            __syncthreads();
        }
        // right-hand side of the diagonal block, and the block itself
        for (int e = tid; e < FB * FB; e += 256) {
            const int i = e >> 6, j = e & 63;
            tile[i][j] = (i < bs && j < bs) ? M[(long)(b0 + i) * lda + b0 + j] : T(0);
        }
#pragma unroll
        for (int q = 0; q < FMAXR / 4; ++q) {
            const int r = wave + 4 * q;
            if (r < nrhs && lane < bs) xs[r][lane] = X[(long)r * ldx + b0 + lane] - acc[q];
        }
        __syncthreads();
        {
#pragma clang fp contract(off)
            for (int r = wave; r < nrhs; r += 4) {
                T x = lane < bs ? xs[r][lane] : T(0);
                if (fwd) {
                    for (int c = 0; c < bs; ++c) {
                        if (MODE == LOWER && lane == c) x = x / tile[c][c];
                        const T xc = __shfl(x, c, 64);
                        if (lane > c && lane < bs) x = x - tile[lane][c] * xc;
                    }
                } else {
                    for (int c = bs - 1; c >= 0; --c) {
                        if (lane == c) x = x / tile[c][c];
                        const T xc = __shfl(x, c, 64);
                        if (lane < c) x = MODE == UPPER ? x - tile[lane][c] * xc : x - xc * tile[c][lane];
                    }
                }
                if (lane < bs) X[(long)r * ldx + b0 + lane] = x;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(&flags[blk], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <class T, int MODE>
static int trsv_flag(hipStream_t s, size_t n, const T *m, size_t lda, size_t nrhs, T *x, size_t ldx) {
    if (n == 0 || nrhs == 0) return RLG_OK;
    const size_t nb = (n + FB - 1) / FB;
    PoolBuf ctl;
    RLG_TRY(ctl.alloc((nb + 2) * 4, s));
    unsigned *flags = ctl.as<unsigned>();
    for (size_t r0 = 0; r0 < nrhs; r0 += FMAXR) {
        const size_t nr = nrhs - r0 < (size_t)FMAXR ? nrhs - r0 : FMAXR;
        RLG_TRY(fill_u32(flags, nb + 2, 0u, s));  // flags, ticket, timeout
        const unsigned grid = (unsigned)(nb < 512 ? nb : 512);
        hipLaunchKernelGGL((k_trsv_flag<T, MODE>), dim3(grid), dim3(256), 0, s, m, (long)lda, (int)n,
                           (int)nr, x + r0 * ldx, (long)ldx, flags + nb, flags,
                           reinterpret_cast<int *>(flags + nb + 1));
        RLG_LAUNCHED();
    }
    return RLG_OK;
}

// Minimum order for the single-launch flag-ordered solve.  The path is off
// (no order reaches it): a solve through it stalled on the GPU and the stall
// has not been traced to its cause, so every solve uses the blocked launches.
constexpr size_t FLAG_MIN_N = ~(size_t)0;

template <class T>
int trsv_lower(hipStream_t s, size_t n, const T *l, size_t lda, bool unit, size_t nrhs, T *x,
               size_t ldx) {
    RLG_PROF("trsv", s);
    if (n >= FLAG_MIN_N)
        return unit ? trsv_flag<T, LOWER_UNIT>(s, n, l, lda, nrhs, x, ldx)
                    : trsv_flag<T, LOWER>(s, n, l, lda, nrhs, x, ldx);
    for (size_t b0 = 0; b0 < n; b0 += TB) {
        const size_t bs = n - b0 < (size_t)TB ? n - b0 : TB;
        if (unit)
            RLG_TRY((diag<T, LOWER_UNIT>(s, l, lda, b0, bs, nrhs, x, ldx)));
        else
            RLG_TRY((diag<T, LOWER>(s, l, lda, b0, bs, nrhs, x, ldx)));
        const size_t r0 = b0 + bs;
        if (r0 < n)  // X^T[:, r0:n] -= X^T[:, b0:r0] * L[r0:n, b0:r0]^T
            RLG_TRY(gemm_t<T>(s, nrhs, n - r0, bs, T(-1), x + b0, ldx, l + r0 * lda + b0, lda, T(1),
                              x + r0, ldx, true, false));
    }
    return RLG_OK;
}

template <class T>
int trsv_upper(hipStream_t s, size_t n, const T *u, size_t lda, size_t nrhs, T *x, size_t ldx) {
    RLG_PROF("trsv", s);
    if (n == 0) return RLG_OK;
    if (n >= FLAG_MIN_N) return trsv_flag<T, UPPER>(s, n, u, lda, nrhs, x, ldx);
    const size_t nb = (n + TB - 1) / TB;
    for (size_t bi = nb; bi-- > 0;) {
        const size_t b0 = bi * TB;
        const size_t bs = n - b0 < (size_t)TB ? n - b0 : TB;
        RLG_TRY((diag<T, UPPER>(s, u, lda, b0, bs, nrhs, x, ldx)));
        if (b0 > 0)  // X^T[:, 0:b0] -= X^T[:, b0:b0+bs] * U[0:b0, b0:b0+bs]^T
            RLG_TRY(gemm_t<T>(s, nrhs, b0, bs, T(-1), x + b0, ldx, u + b0, lda, T(1), x, ldx, true,
                              false));
    }
    return RLG_OK;
}

template <class T>
int trsv_lower_t(hipStream_t s, size_t n, const T *l, size_t lda, size_t nrhs, T *x, size_t ldx) {
    RLG_PROF("trsv", s);
    if (n == 0) return RLG_OK;
    if (n >= FLAG_MIN_N) return trsv_flag<T, LOWER_T>(s, n, l, lda, nrhs, x, ldx);
    const size_t nb = (n + TB - 1) / TB;
    for (size_t bi = nb; bi-- > 0;) {
        const size_t b0 = bi * TB;
        const size_t bs = n - b0 < (size_t)TB ? n - b0 : TB;
        RLG_TRY((diag<T, LOWER_T>(s, l, lda, b0, bs, nrhs, x, ldx)));
        if (b0 > 0)  // X^T[:, 0:b0] -= X^T[:, b0:b0+bs] * L[b0:b0+bs, 0:b0]
            RLG_TRY(gemm_t<T>(s, nrhs, b0, bs, T(-1), x + b0, ldx, l + b0 * lda, lda, T(1), x, ldx,
                              false, false));
    }
    return RLG_OK;
}

template int trsv_lower<double>(hipStream_t, size_t, const double *, size_t, bool, size_t, double *,
                                size_t);
template int trsv_lower<float>(hipStream_t, size_t, const float *, size_t, bool, size_t, float *,
                               size_t);
template int trsv_upper<double>(hipStream_t, size_t, const double *, size_t, size_t, double *,
                                size_t);
template int trsv_upper<float>(hipStream_t, size_t, const float *, size_t, size_t, float *, size_t);
template int trsv_lower_t<double>(hipStream_t, size_t, const double *, size_t, size_t, double *,
                                  size_t);
template int trsv_lower_t<float>(hipStream_t, size_t, const float *, size_t, size_t, float *,
                                 size_t);

}  // namespace rlg
