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

template <class T>
int trsv_lower(hipStream_t s, size_t n, const T *l, size_t lda, bool unit, size_t nrhs, T *x,
               size_t ldx) {
    RLG_PROF("trsv", s);
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
