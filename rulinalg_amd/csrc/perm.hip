// PermutationMatrix * Matrix and Matrix * PermutationMatrix on the device
// (src/matrix/impl_permutation_mul.rs:53-204): pure data movement, bound by
// HBM.  The reference's into-buffer forms define the result
// (src/matrix/permutation_matrix.rs:318-358): row i of the source becomes row
// perm[i] of the result (P * M), column j becomes column perm[j] (M * P).
// The in-place forms of the reference (permute_by_swap, permutation_matrix.rs:
// 479-509) produce the same matrix.
#include <vector>

#include "kernels.hpp"

namespace rlg {

// out[perm[i]][c] = a[i][c]: one workgroup moves 8 rows, each row by a
// wave with contiguous (coalesced) reads and writes.
template <class T>
__global__ __launch_bounds__(256) void k_permute_rows(int rows, int cols, const int64_t *__restrict__ perm,
                                                      const T *__restrict__ a, long lda, T *__restrict__ out,
                                                      long ldo) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int r = blockIdx.x * 8 + wave * 2; r < min(rows, (int)blockIdx.x * 8 + wave * 2 + 2); ++r) {
        const long dst = perm[r];
        const T *src = a + (long)r * lda;
        T *d = out + dst * ldo;
        for (int c = lane + (int)blockIdx.y * 64 * 32; c < min(cols, ((int)blockIdx.y + 1) * 64 * 32); c += 64)
            d[c] = src[c];
    }
}

// out[i][perm[j]] = a[i][j], as out[i][q] = a[i][inv[q]] so that the writes
// are contiguous; the reads of a row stay within the row (one cache-resident
// line set per row).
template <class T>
__global__ __launch_bounds__(256) void k_permute_cols(int rows, int cols, const int64_t *__restrict__ inv,
                                                      const T *__restrict__ a, long lda, T *__restrict__ out,
                                                      long ldo) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= cols) return;
    const long src = inv[q];
    for (int i = blockIdx.y; i < rows; i += gridDim.y) out[(long)i * ldo + q] = a[(long)i * lda + src];
}

__global__ void k_invert_perm(int n, const int64_t *__restrict__ perm, int64_t *__restrict__ inv) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) inv[perm[i]] = i;
}

template <class T>
int permute_rows_dev(hipStream_t s, size_t n, const int64_t *perm, size_t cols, const T *a, size_t lda, T *out,
                     size_t ldo) {
    if (n == 0 || cols == 0) return RLG_OK;
    RLG_PROF_W("permute", s, 2.0 * (double)n * (double)cols * sizeof(T));
    hipLaunchKernelGGL(k_permute_rows<T>, dim3(ceil_div(n, 8), ceil_div(cols, 64 * 32)), dim3(256), 0, s,
                       (int)n, (int)cols, perm, a, (long)lda, out, (long)ldo);
    RLG_LAUNCHED();
    return RLG_OK;
}

template <class T>
int permute_cols_dev(hipStream_t s, size_t rows, size_t n, const int64_t *perm, const T *a, size_t lda, T *out,
                     size_t ldo, int64_t *inv) {
    if (n == 0 || rows == 0) return RLG_OK;
    RLG_PROF_W("permute", s, 2.0 * (double)n * (double)rows * sizeof(T));
    hipLaunchKernelGGL(k_invert_perm, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n, perm, inv);
    RLG_LAUNCHED();
    const unsigned gy = (unsigned)(rows < 1024 ? rows : 1024);
    hipLaunchKernelGGL(k_permute_cols<T>, dim3(ceil_div(n, 256), gy), dim3(256), 0, s, (int)rows, (int)n, inv, a,
                       (long)lda, out, (long)ldo);
    RLG_LAUNCHED();
    return RLG_OK;
}

// validate_permutation (permutation_matrix.rs:438-467): every entry in
// [0, n) and no duplicates.
static int validate(size_t n, const size_t *perm) {
    std::vector<char> seen(n, 0);
    for (size_t i = 0; i < n; ++i) {
        if (perm[i] >= n) {
            set_error("Supplied permutation array contains elements out of bounds.");
            return RLG_E_INVALID_PERMUTATION;
        }
        seen[perm[i]] = 1;
    }
    for (size_t i = 0; i < n; ++i)
        if (!seen[i]) {
            set_error("Supplied permutation array contains duplicate elements.");
            return RLG_E_INVALID_PERMUTATION;
        }
    return RLG_OK;
}

template <class T>
static int permute_host(bool by_rows, size_t perm_len, const size_t *perm, size_t rows, size_t cols, const T *a,
                        size_t lda, T *out, size_t ldo) {
    if (by_rows ? perm_len != rows : perm_len != cols) {
        set_error(by_rows ? "Permutation matrix and right-hand side matrix dimensions are not compatible."
                          : "Left-hand side matrix and permutation matrix dimensions are not compatible.");
        return RLG_E_SHAPE;
    }
    if (lda < cols || ldo < cols) { set_error("row stride smaller than the column count"); return RLG_E_INVALID_ARG; }
    if (perm_len && !perm) { set_error("null permutation"); return RLG_E_INVALID_ARG; }
    RLG_TRY(validate(perm_len, perm));
    if (rows == 0 || cols == 0) return RLG_OK;
    if (!a || !out) { set_error("null operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    PoolBuf da, dout, dp, dinv;
    RLG_TRY(da.alloc(rows * cols * sizeof(T), s));
    RLG_TRY(dout.alloc(rows * cols * sizeof(T), s));
    RLG_TRY(dp.alloc(perm_len * 8, s));
    RLG_TRY(dinv.alloc(perm_len * 8, s));
    std::vector<int64_t> p64(perm, perm + perm_len);
    RLG_HIP(hipMemcpy2DAsync(da.p, cols * sizeof(T), a, lda * sizeof(T), cols * sizeof(T), rows,
                             hipMemcpyHostToDevice, s));
    RLG_HIP(hipMemcpyAsync(dp.p, p64.data(), perm_len * 8, hipMemcpyHostToDevice, s));
    if (by_rows)
        RLG_TRY(permute_rows_dev<T>(s, rows, dp.as<int64_t>(), cols, da.as<T>(), cols, dout.as<T>(), cols));
    else
        RLG_TRY(permute_cols_dev<T>(s, rows, cols, dp.as<int64_t>(), da.as<T>(), cols, dout.as<T>(), cols,
                                    dinv.as<int64_t>()));
    RLG_HIP(hipMemcpy2DAsync(out, ldo * sizeof(T), dout.p, cols * sizeof(T), cols * sizeof(T), rows,
                             hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

}  // namespace rlg

using namespace rlg;

extern "C" {

int rlg_permute_rows_f64(size_t perm_len, const size_t *perm, size_t rows, size_t cols, const double *a,
                         size_t lda, double *out, size_t ldo) {
    return permute_host<double>(true, perm_len, perm, rows, cols, a, lda, out, ldo);
}
int rlg_permute_rows_f32(size_t perm_len, const size_t *perm, size_t rows, size_t cols, const float *a,
                         size_t lda, float *out, size_t ldo) {
    return permute_host<float>(true, perm_len, perm, rows, cols, a, lda, out, ldo);
}
int rlg_permute_cols_f64(size_t perm_len, const size_t *perm, size_t rows, size_t cols, const double *a,
                         size_t lda, double *out, size_t ldo) {
    return permute_host<double>(false, perm_len, perm, rows, cols, a, lda, out, ldo);
}
int rlg_permute_cols_f32(size_t perm_len, const size_t *perm, size_t rows, size_t cols, const float *a,
                         size_t lda, float *out, size_t ldo) {
    return permute_host<float>(false, perm_len, perm, rows, cols, a, lda, out, ldo);
}

int rlg_permute_rows_f64_dev(size_t n, const int64_t *perm, size_t cols, const double *a, size_t lda,
                             double *out, size_t ldo, void *stream) {
    if (lda < cols || ldo < cols || (n && (!perm || !a || !out))) {
        set_error("bad operand");
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    return permute_rows_dev<double>(pick_stream(d, stream), n, perm, cols, a, lda, out, ldo);
}

int rlg_permute_cols_f64_dev(size_t rows, size_t n, const int64_t *perm, const double *a, size_t lda,
                             double *out, size_t ldo, void *stream) {
    if (lda < n || ldo < n || (n && rows && (!perm || !a || !out))) {
        set_error("bad operand");
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = pick_stream(d, stream);
    PoolBuf inv;  // back to the pool behind an event: the call stays asynchronous
    RLG_TRY(inv.alloc(n * 8, s));
    return permute_cols_dev<double>(s, rows, n, perm, a, lda, out, ldo, inv.as<int64_t>());
}

}  // extern "C"
