// The C ABI of librlgpu for GEMM, Mat*Vec, PartialPivLu and Cholesky.
// Host-pointer entry points mirror the reference's synchronous calls
// (copy in, compute on the device, copy out); _dev entry points work on
// device-resident operands and leave the data on the device.
#include <cmath>
#include <vector>

#include "kernels.hpp"

using namespace rlg;

// Factor handles: immutable after decompose, so concurrent solves are safe.
struct rlg_lu {
    int elem = 8;           // 8: f64 factor, 4: f32 factor
    size_t n = 0, lda = 0;
    void *a = nullptr;      // packed factor on the device
    PoolBuf own;            // backing store when the handle owns `a`
    PoolBuf raw_d;          // raw permutation (int64) on the device
    std::vector<int64_t> ipiv, raw;
    mutable StreamUse use;  // streams of asynchronous solves on this factor
};

struct rlg_chol {
    int elem = 8;
    size_t n = 0, lda = 0;
    void *a = nullptr;
    PoolBuf own;
    mutable StreamUse use;
};

namespace {

template <class T>
int upload_matrix(hipStream_t s, T *dst, size_t ldd, const T *src, size_t lds, size_t rows,
                  size_t cols) {
    if (rows == 0 || cols == 0) return RLG_OK;
    RLG_HIP(hipMemcpy2DAsync(dst, ldd * sizeof(T), src, lds * sizeof(T), cols * sizeof(T), rows,
                             hipMemcpyHostToDevice, s));
    return RLG_OK;
}

template <class T>
int download_matrix(hipStream_t s, T *dst, size_t ldd, const T *src, size_t lds, size_t rows,
                    size_t cols) {
    if (rows == 0 || cols == 0) return RLG_OK;
    RLG_HIP(hipMemcpy2DAsync(dst, ldd * sizeof(T), src, lds * sizeof(T), cols * sizeof(T), rows,
                             hipMemcpyDeviceToHost, s));
    return RLG_OK;
}

int check_out(const void *p) {
    if (!p) { set_error("null output pointer"); return RLG_E_INVALID_ARG; }
    return RLG_OK;
}

// ---------------------------------------------------------------- GEMM

template <class T>
int gemm_host(size_t a_rows, size_t a_cols, const T *a, size_t lda, size_t b_rows, size_t b_cols,
              const T *b, size_t ldb, T *c, size_t ldc) {
    if (a_cols != b_rows) {
        set_error("Matrix dimensions do not agree: %zu x %zu * %zu x %zu", a_rows, a_cols, b_rows,
                  b_cols);
        return RLG_E_SHAPE;
    }
    if (lda < a_cols || ldb < b_cols || ldc < b_cols) {
        set_error("row stride smaller than the column count");
        return RLG_E_INVALID_ARG;
    }
    if (a_rows == 0 || b_cols == 0) return RLG_OK;
    if ((!a || !b) && a_cols) { set_error("null operand"); return RLG_E_INVALID_ARG; }
    RLG_TRY(check_out(c));
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    const size_t m = a_rows, k = a_cols, n = b_cols;
    DevBuf da, db, dc;
    RLG_TRY(da.alloc(m * k * sizeof(T)));
    RLG_TRY(db.alloc(k * n * sizeof(T)));
    RLG_TRY(dc.alloc(m * n * sizeof(T)));
    RLG_TRY(upload_matrix<T>(s, da.as<T>(), k, a, lda, m, k));
    RLG_TRY(upload_matrix<T>(s, db.as<T>(), n, b, ldb, k, n));
    if (k == 0) {
        // empty inner dimension: the product is all zeros (mat_mul.rs:76-98 starts from zero)
        std::vector<T> z(n, T(0));
        for (size_t i = 0; i < m; ++i) std::memcpy(c + i * ldc, z.data(), n * sizeof(T));
        return RLG_OK;
    }
    RLG_TRY(gemm_t<T>(s, m, n, k, T(1), da.as<T>(), k, db.as<T>(), n, T(0), dc.as<T>(), n));
    RLG_TRY(download_matrix<T>(s, c, ldc, dc.as<T>(), n, m, n));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

// ---------------------------------------------------------------- LU

template <class T>
int lu_finish(hipStream_t s, rlg_lu *h, int64_t *ipiv_d, int64_t *info_d, void *work) {
    const size_t n = h->n;
    int64_t info = 0;
    int timeout = 0;
    h->ipiv.assign(n, 0);
    h->raw.assign(n, 0);
    if (n) RLG_HIP(hipMemcpyAsync(h->ipiv.data(), ipiv_d, n * 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipMemcpyAsync(&info, info_d, 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipMemcpyAsync(&timeout, lu_timeout_flag(work, n, sizeof(T)), 4, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    if (timeout) {
        set_error("LU panel barrier timed out (workgroups not co-resident)");
        return RLG_E_DEVICE;
    }
    if (info < (int64_t)n) {
        set_error("The matrix is too ill-conditioned for LU decomposition with partial pivoting "
                  "(pivot %lld below epsilon)", (long long)info);
        return RLG_E_DIV_BY_ZERO;
    }
    perm_from_ipiv(n, h->ipiv.data(), h->raw.data());
    RLG_TRY(h->raw_d.alloc(n * 8, s));
    if (n) RLG_HIP(hipMemcpyAsync(h->raw_d.p, h->raw.data(), n * 8, hipMemcpyHostToDevice, s));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

template <class T>
int lu_factor_handle(hipStream_t s, rlg_lu *h) {
    const size_t n = h->n;
    PoolBuf ipiv, info, work;
    RLG_TRY(ipiv.alloc(n * 8, s));
    RLG_TRY(info.alloc(8, s));
    RLG_TRY(work.alloc(lu_work_bytes(n, sizeof(T)), s));
    RLG_TRY(lu_factor<T>(s, n, static_cast<T *>(h->a), h->lda, ipiv.as<int64_t>(), info.as<int64_t>(),
                         work.p));
    return lu_finish<T>(s, h, ipiv.as<int64_t>(), info.as<int64_t>(), work.p);
}

template <class T>
int lu_decompose_host(size_t rows, size_t cols, const T *a, size_t lda, rlg_lu **out) {
    RLG_TRY(check_out(out));
    *out = nullptr;
    if (rows != cols) {
        set_error("Matrix must be square for LU decomposition (%zu x %zu).", rows, cols);
        return RLG_E_SHAPE;
    }
    if (lda < cols || (!a && rows)) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    auto *h = new rlg_lu();
    h->elem = sizeof(T);
    h->n = rows;
    h->lda = rows;
    int st = h->own.alloc(rows * rows * sizeof(T), s);
    if (st == RLG_OK) {
        h->a = h->own.p;
        st = upload_matrix<T>(s, static_cast<T *>(h->a), rows, a, lda, rows, rows);
    }
    if (st == RLG_OK) st = lu_factor_handle<T>(s, h);
    if (st != RLG_OK) { delete h; return st; }
    *out = h;
    return RLG_OK;
}

template <class T>
int lu_solve_host(const rlg_lu *h, size_t b_len, size_t nrhs, const T *b, size_t ldb, T *x,
                  size_t ldx) {
    if (!h || h->elem != (int)sizeof(T)) { set_error("factor handle of another precision"); return RLG_E_INVALID_ARG; }
    const size_t n = h->n;
    if (b_len != n) {
        set_error("Right-hand side vector must have compatible size (%zu vs %zu).", b_len, n);
        return RLG_E_SHAPE;
    }
    if (nrhs == 0 || n == 0) return RLG_OK;
    if (!b || !x || ldb < n || ldx < n) { set_error("bad right-hand side"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    PoolBuf db, tmp;
    RLG_TRY(db.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(tmp.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(upload_matrix<T>(s, db.as<T>(), n, b, ldb, nrhs, n));
    RLG_TRY(lu_solve<T>(s, n, static_cast<const T *>(h->a), h->lda, h->raw_d.as<int64_t>(), nrhs,
                        db.as<T>(), n, db.as<T>(), n, tmp.as<T>()));
    RLG_TRY(download_matrix<T>(s, x, ldx, db.as<T>(), n, nrhs, n));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

// ---------------------------------------------------------------- Cholesky

template <class T>
int chol_factor_handle(hipStream_t s, rlg_chol *h) {
    const size_t n = h->n;
    PoolBuf info, work;
    RLG_TRY(info.alloc(8, s));
    RLG_TRY(work.alloc(chol_work_bytes(n, sizeof(T)), s));
    RLG_TRY(chol_factor<T>(s, n, static_cast<T *>(h->a), h->lda, info.as<int64_t>(), work.p));
    int64_t inf = 0;
    RLG_HIP(hipMemcpyAsync(&inf, info.p, 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    if (inf < (int64_t)n) {
        set_error("Cholesky: diagonal entry %lld is not positive to working precision", (long long)inf);
        return RLG_E_DECOMP_FAILURE;
    }
    return RLG_OK;
}

template <class T>
int chol_decompose_host(size_t rows, size_t cols, const T *a, size_t lda, rlg_chol **out) {
    RLG_TRY(check_out(out));
    *out = nullptr;
    if (rows != cols) {
        set_error("Matrix must be square for Cholesky decomposition (%zu x %zu).", rows, cols);
        return RLG_E_SHAPE;
    }
    if (lda < cols || (!a && rows)) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    auto *h = new rlg_chol();
    h->elem = sizeof(T);
    h->n = rows;
    h->lda = rows;
    int st = h->own.alloc(rows * rows * sizeof(T), s);
    if (st == RLG_OK) {
        h->a = h->own.p;
        st = upload_matrix<T>(s, static_cast<T *>(h->a), rows, a, lda, rows, rows);
    }
    if (st == RLG_OK) st = chol_factor_handle<T>(s, h);
    if (st != RLG_OK) { delete h; return st; }
    *out = h;
    return RLG_OK;
}

template <class T>
int chol_solve_host(const rlg_chol *h, size_t b_len, size_t nrhs, const T *b, size_t ldb, T *x,
                    size_t ldx) {
    if (!h || h->elem != (int)sizeof(T)) { set_error("factor handle of another precision"); return RLG_E_INVALID_ARG; }
    const size_t n = h->n;
    if (b_len != n) {
        set_error("RHS vector and coefficient matrix must be dimensionally compatible (%zu vs %zu).",
                  b_len, n);
        return RLG_E_SHAPE;
    }
    if (nrhs == 0 || n == 0) return RLG_OK;
    if (!b || !x || ldb < n || ldx < n) { set_error("bad right-hand side"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    PoolBuf db, tmp;
    RLG_TRY(db.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(tmp.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(upload_matrix<T>(s, db.as<T>(), n, b, ldb, nrhs, n));
    RLG_TRY(chol_solve<T>(s, n, static_cast<const T *>(h->a), h->lda, nrhs, db.as<T>(), n, db.as<T>(), n,
                          tmp.as<T>()));
    RLG_TRY(download_matrix<T>(s, x, ldx, db.as<T>(), n, nrhs, n));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

template <class T>
int chol_unpack(const rlg_chol *h, T *l, size_t ld) {
    if (!h || h->elem != (int)sizeof(T)) { set_error("factor handle of another precision"); return RLG_E_INVALID_ARG; }
    RLG_TRY(check_out(l));
    const size_t n = h->n;
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_TRY(download_matrix<T>(d->stream, l, ld, static_cast<const T *>(h->a), h->lda, n, n));
    RLG_HIP(hipStreamSynchronize(d->stream));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) l[i * ld + j] = T(0);  // nullify_upper_triangular_part
    return RLG_OK;
}

}  // namespace

extern "C" {

int rlg_gemm_f64(size_t a_rows, size_t a_cols, const double *a, size_t lda, size_t b_rows,
                 size_t b_cols, const double *b, size_t ldb, double *c, size_t ldc) {
    return gemm_host<double>(a_rows, a_cols, a, lda, b_rows, b_cols, b, ldb, c, ldc);
}

int rlg_gemm_f32(size_t a_rows, size_t a_cols, const float *a, size_t lda, size_t b_rows,
                 size_t b_cols, const float *b, size_t ldb, float *c, size_t ldc) {
    return gemm_host<float>(a_rows, a_cols, a, lda, b_rows, b_cols, b, ldb, c, ldc);
}

int rlg_gemm_f64_dev(size_t m, size_t k, size_t n, double alpha, const double *a, size_t lda,
                     const double *b, size_t ldb, double beta, double *c, size_t ldc, void *stream) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (lda < k || ldb < n || ldc < n) { set_error("row stride smaller than the column count"); return RLG_E_INVALID_ARG; }
    return gemm_f64(pick_stream(d, stream), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int rlg_gemm_f32_dev(size_t m, size_t k, size_t n, float alpha, const float *a, size_t lda,
                     const float *b, size_t ldb, float beta, float *c, size_t ldc, void *stream) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (lda < k || ldb < n || ldc < n) { set_error("row stride smaller than the column count"); return RLG_E_INVALID_ARG; }
    return gemm_f32(pick_stream(d, stream), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int rlg_gemm_f32_split_dev(size_t m, size_t k, size_t n, float alpha, const float *a, size_t lda,
                           const float *b, size_t ldb, float beta, float *c, size_t ldc, void *stream) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (lda < k || ldb < n || ldc < n) { set_error("row stride smaller than the column count"); return RLG_E_INVALID_ARG; }
    if (m == 0 || n == 0) return RLG_OK;
    if (!a || !b || !c) { set_error("null operand"); return RLG_E_INVALID_ARG; }
    hipStream_t s = pick_stream(d, stream);
    if (k == 0) return gemm_f32(s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);  // C = beta C
    const size_t bytes = gemm_lowp_pack_bytes(m, n, k);
    PoolBuf pack;
    RLG_TRY(pack.alloc(bytes, s));
    return gemm_lowp(s, RLG_PREC_F16X3, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, false, false, pack.p,
                     bytes);
}

int rlg_matvec_f64(size_t m, size_t n, const double *a, size_t lda, size_t v_len, const double *v,
                   double *y) {
    if (v_len != n) {
        set_error("Matrix and Vector dimensions do not agree (%zu vs %zu).", n, v_len);
        return RLG_E_SHAPE;
    }
    if (m == 0) return RLG_OK;
    RLG_TRY(check_out(y));
    if (n == 0) { for (size_t i = 0; i < m; ++i) y[i] = 0.0; return RLG_OK; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    DevBuf da, dv, dy;
    RLG_TRY(da.alloc(m * n * 8));
    RLG_TRY(dv.alloc(n * 8));
    RLG_TRY(dy.alloc(m * 8));
    RLG_TRY(upload_matrix<double>(s, da.as<double>(), n, a, lda, m, n));
    RLG_HIP(hipMemcpyAsync(dv.p, v, n * 8, hipMemcpyHostToDevice, s));
    RLG_TRY(gemv_f64(s, m, n, da.as<double>(), n, dv.as<double>(), dy.as<double>()));
    RLG_HIP(hipMemcpyAsync(y, dy.p, m * 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

// ------------------------------------------------------------------ triangular

}  // extern "C"

namespace {

// solve_u_triangular / solve_l_triangular (src/matrix/base/mod.rs:1015-1024,
// 1058-1067): the vector must match the column count (panic), the
// substitution needs a square matrix (panic), and a diagonal entry below
// epsilon in magnitude fails with DivByZero (src/matrix/mod.rs:331-335,
// 375-378).  Only the triangle named by `upper` is read.
template <class T>
int tri_solve_dev(hipStream_t s, bool upper, size_t n, const T *m, size_t ldm, size_t nrhs,
                  const T *b, size_t ldb, T *x, size_t ldx) {
    PoolBuf flag, tmp;
    RLG_TRY(flag.alloc(4, s));
    RLG_TRY(tmp.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(fill_u32(flag.as<uint32_t>(), 1, 0u, s));
    RLG_TRY(diag_tiny<T>(s, n, m, ldm, std::numeric_limits<T>::epsilon(), flag.as<uint32_t>()));
    // b goes to tmp first (x may alias b); tmp is the running residual
    RLG_HIP(hipMemcpy2DAsync(tmp.p, n * sizeof(T), b, ldb * sizeof(T), n * sizeof(T), nrhs,
                             hipMemcpyDeviceToDevice, s));
    if (upper)
        RLG_TRY(trsv_upper<T>(s, n, m, ldm, nrhs, tmp.as<T>(), n, x, ldx));
    else
        RLG_TRY(trsv_lower<T>(s, n, m, ldm, false, nrhs, tmp.as<T>(), n, x, ldx));
    uint32_t tiny = 0;
    RLG_HIP(hipMemcpyAsync(&tiny, flag.p, 4, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    if (tiny) {
        set_error("%s triangular matrix is singular to working precision.", upper ? "Upper" : "Lower");
        return RLG_E_DIV_BY_ZERO;
    }
    return RLG_OK;
}

template <class T>
int tri_solve_host(bool upper, size_t rows, size_t cols, const T *m, size_t ldm, size_t b_len,
                   size_t nrhs, const T *b, size_t ldb, T *x, size_t ldx) {
    if (b_len != cols) {
        set_error("Vector size %zu != %zu Matrix column count.", b_len, cols);
        return RLG_E_SHAPE;
    }
    if (rows != cols) {
        set_error("Matrix %s must be square.", upper ? "U" : "L");
        return RLG_E_SHAPE;
    }
    const size_t n = rows;
    if (n == 0 || nrhs == 0) return RLG_OK;
    if (!m || !b || !x || ldm < n || ldb < n || ldx < n) {
        set_error("bad operand");
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    PoolBuf dm, dx;
    RLG_TRY(dm.alloc(n * n * sizeof(T), s));
    RLG_TRY(dx.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(upload_matrix<T>(s, dm.as<T>(), n, m, ldm, n, n));
    RLG_TRY(upload_matrix<T>(s, dx.as<T>(), n, b, ldb, nrhs, n));
    const int st = tri_solve_dev<T>(s, upper, n, dm.as<T>(), n, nrhs, dx.as<T>(), n, dx.as<T>(), n);
    if (st != RLG_OK) return st;
    RLG_TRY(download_matrix<T>(s, x, ldx, dx.as<T>(), n, nrhs, n));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

}  // namespace

extern "C" {

int rlg_solve_u_triangular_f64(size_t rows, size_t cols, const double *u, size_t ldu, size_t b_len,
                               size_t nrhs, const double *b, size_t ldb, double *x, size_t ldx) {
    return tri_solve_host<double>(true, rows, cols, u, ldu, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_solve_l_triangular_f64(size_t rows, size_t cols, const double *l, size_t ldl, size_t b_len,
                               size_t nrhs, const double *b, size_t ldb, double *x, size_t ldx) {
    return tri_solve_host<double>(false, rows, cols, l, ldl, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_solve_u_triangular_f32(size_t rows, size_t cols, const float *u, size_t ldu, size_t b_len,
                               size_t nrhs, const float *b, size_t ldb, float *x, size_t ldx) {
    return tri_solve_host<float>(true, rows, cols, u, ldu, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_solve_l_triangular_f32(size_t rows, size_t cols, const float *l, size_t ldl, size_t b_len,
                               size_t nrhs, const float *b, size_t ldb, float *x, size_t ldx) {
    return tri_solve_host<float>(false, rows, cols, l, ldl, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_solve_triangular_f64_dev(int upper, size_t n, const double *m, size_t ldm, size_t nrhs,
                                 const double *b, size_t ldb, double *x, size_t ldx, void *stream) {
    if (n == 0 || nrhs == 0) return RLG_OK;
    if (!m || !b || !x || ldm < n || ldb < n || ldx < n) {
        set_error("bad operand");
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    return tri_solve_dev<double>(pick_stream(d, stream), upper != 0, n, m, ldm, nrhs, b, ldb, x, ldx);
}

// ------------------------------------------------------------------ LU

int rlg_lu_decompose_f64(size_t rows, size_t cols, const double *a, size_t lda, rlg_lu **out) {
    return lu_decompose_host<double>(rows, cols, a, lda, out);
}

int rlg_lu_decompose_f32(size_t rows, size_t cols, const float *a, size_t lda, rlg_lu **out) {
    return lu_decompose_host<float>(rows, cols, a, lda, out);
}

int rlg_lu_decompose_f64_dev(size_t n, double *a, size_t lda, void *stream, rlg_lu **out) {
    RLG_TRY(check_out(out));
    *out = nullptr;
    if (lda < n || (!a && n)) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    auto *h = new rlg_lu();
    h->elem = 8;
    h->n = n;
    h->lda = lda;
    h->a = a;
    int st = lu_factor_handle<double>(pick_stream(d, stream), h);
    if (st != RLG_OK) { delete h; return st; }
    *out = h;
    return RLG_OK;
}

int rlg_lu_solve_f64(const rlg_lu *lu, size_t b_len, size_t nrhs, const double *b, size_t ldb,
                     double *x, size_t ldx) {
    return lu_solve_host<double>(lu, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_lu_solve_f32(const rlg_lu *lu, size_t b_len, size_t nrhs, const float *b, size_t ldb,
                     float *x, size_t ldx) {
    return lu_solve_host<float>(lu, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_lu_solve_f64_dev(const rlg_lu *lu, size_t nrhs, const double *b, size_t ldb, double *x,
                         size_t ldx, void *stream) {
    if (!lu || lu->elem != 8) { set_error("not an f64 factor"); return RLG_E_INVALID_ARG; }
    if (nrhs == 0 || lu->n == 0) return RLG_OK;
    if (ldb < lu->n || ldx < lu->n) { set_error("bad right-hand side stride"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = pick_stream(d, stream);
    PoolBuf tmp;  // returned to the pool with an event: the call stays asynchronous
    RLG_TRY(tmp.alloc(nrhs * lu->n * 8, s));
    lu->use.add(s);
    return lu_solve<double>(s, lu->n, static_cast<const double *>(lu->a), lu->lda,
                            lu->raw_d.as<int64_t>(), nrhs, b, ldb, x, ldx, tmp.as<double>());
}

int rlg_lu_size(const rlg_lu *lu, size_t *n) {
    if (!lu || !n) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    *n = lu->n;
    return RLG_OK;
}

int rlg_lu_perm(const rlg_lu *lu, size_t *perm) {
    if (!lu || (!perm && lu->n)) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    // stored p = inverse of the swap product: perm[raw[k]] = k
    for (size_t k = 0; k < lu->n; ++k) perm[lu->raw[k]] = k;
    return RLG_OK;
}

int rlg_lu_ipiv(const rlg_lu *lu, int64_t *ipiv) {
    if (!lu || (!ipiv && lu->n)) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    for (size_t k = 0; k < lu->n; ++k) ipiv[k] = lu->ipiv[k];
    return RLG_OK;
}

int rlg_lu_packed_f64(const rlg_lu *lu, double *out, size_t ld) {
    if (!lu || lu->elem != 8) { set_error("not an f64 factor"); return RLG_E_INVALID_ARG; }
    RLG_TRY(check_out(out));
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_HIP(hipDeviceSynchronize());
    RLG_TRY(download_matrix<double>(d->stream, out, ld, static_cast<const double *>(lu->a), lu->lda,
                                    lu->n, lu->n));
    RLG_HIP(hipStreamSynchronize(d->stream));
    return RLG_OK;
}

int rlg_lu_packed_f32(const rlg_lu *lu, float *out, size_t ld) {
    if (!lu || lu->elem != 4) { set_error("not an f32 factor"); return RLG_E_INVALID_ARG; }
    RLG_TRY(check_out(out));
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_TRY(download_matrix<float>(d->stream, out, ld, static_cast<const float *>(lu->a), lu->lda,
                                   lu->n, lu->n));
    RLG_HIP(hipStreamSynchronize(d->stream));
    return RLG_OK;
}

int rlg_lu_unpack_f64(const rlg_lu *lu, double *l, double *u, size_t *perm) {
    if (!lu || lu->elem != 8) { set_error("not an f64 factor"); return RLG_E_INVALID_ARG; }
    const size_t n = lu->n;
    if ((!l || !u || !perm) && n) { set_error("null output"); return RLG_E_INVALID_ARG; }
    std::vector<double> packed(n * n);
    RLG_TRY(rlg_lu_packed_f64(lu, packed.data(), n));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            const double v = packed[i * n + j];
            l[i * n + j] = j < i ? v : (j == i ? 1.0 : 0.0);  // unit_lower_triangular_part
            u[i * n + j] = j < i ? 0.0 : v;                    // nullify_lower_triangular_part
        }
    return rlg_lu_perm(lu, perm);
}

int rlg_lu_det_f64(const rlg_lu *lu, double *det) {
    if (!lu || lu->elem != 8 || !det) { set_error("bad argument"); return RLG_E_INVALID_ARG; }
    const size_t n = lu->n;
    std::vector<double> diag(n);
    Device *d;
    RLG_TRY(current_device(&d));
    // the diagonal as one strided copy (row pitch lda + 1)
    if (n)
        RLG_HIP(hipMemcpy2DAsync(diag.data(), 8, lu->a, (lu->lda + 1) * 8, 8, n,
                                 hipMemcpyDeviceToHost, d->stream));
    RLG_HIP(hipStreamSynchronize(d->stream));
    // lu.rs:291-300: fold the diagonal from one, times the parity of p.
    double u_det = 1.0;
    for (size_t i = 0; i < n; ++i) u_det = u_det * diag[i];
    // parity of the stored permutation equals the parity of the swap count
    size_t swaps = 0;
    for (size_t k = 0; k < n; ++k) swaps += (size_t)lu->ipiv[k] != k;
    const double p_det = (swaps & 1) ? 0.0 - 1.0 : 1.0;
    *det = p_det * u_det;
    return RLG_OK;
}

int rlg_lu_inverse_f64(const rlg_lu *lu, double *inv, size_t ld) {
    if (!lu || lu->elem != 8 || (!inv && lu->n)) { set_error("bad argument"); return RLG_E_INVALID_ARG; }
    const size_t n = lu->n;
    if (n == 0) return RLG_OK;
    // lu.rs:251-285: solve against e_i and store the result as column i.
    std::vector<double> e(n * n, 0.0), x(n * n);
    for (size_t i = 0; i < n; ++i) e[i * n + i] = 1.0;
    RLG_TRY(rlg_lu_solve_f64(lu, n, n, e.data(), n, x.data(), n));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) inv[j * ld + i] = x[i * n + j];
    return RLG_OK;
}

void rlg_lu_free(rlg_lu *lu) {
    if (!lu) return;
    // asynchronous _dev solves may still read the factor and permutation: the
    // memory goes back to the pool behind events on the streams they ran on
    const auto streams = lu->use.all();
    lu->own.release_after(streams);
    lu->raw_d.release_after(streams);
    delete lu;
}

// ------------------------------------------------------------------ Cholesky

int rlg_chol_decompose_f64(size_t rows, size_t cols, const double *a, size_t lda, rlg_chol **out) {
    return chol_decompose_host<double>(rows, cols, a, lda, out);
}

int rlg_chol_decompose_f32(size_t rows, size_t cols, const float *a, size_t lda, rlg_chol **out) {
    return chol_decompose_host<float>(rows, cols, a, lda, out);
}

int rlg_chol_decompose_f64_dev(size_t n, double *a, size_t lda, void *stream, rlg_chol **out) {
    RLG_TRY(check_out(out));
    *out = nullptr;
    if (lda < n || (!a && n)) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    auto *h = new rlg_chol();
    h->elem = 8;
    h->n = n;
    h->lda = lda;
    h->a = a;
    int st = chol_factor_handle<double>(pick_stream(d, stream), h);
    if (st != RLG_OK) { delete h; return st; }
    *out = h;
    return RLG_OK;
}

int rlg_chol_solve_f64(const rlg_chol *ch, size_t b_len, size_t nrhs, const double *b, size_t ldb,
                       double *x, size_t ldx) {
    return chol_solve_host<double>(ch, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_chol_solve_f32(const rlg_chol *ch, size_t b_len, size_t nrhs, const float *b, size_t ldb,
                       float *x, size_t ldx) {
    return chol_solve_host<float>(ch, b_len, nrhs, b, ldb, x, ldx);
}

int rlg_chol_solve_f64_dev(const rlg_chol *ch, size_t nrhs, const double *b, size_t ldb, double *x,
                           size_t ldx, void *stream) {
    if (!ch || ch->elem != 8) { set_error("not an f64 factor"); return RLG_E_INVALID_ARG; }
    if (nrhs == 0 || ch->n == 0) return RLG_OK;
    if (ldb < ch->n || ldx < ch->n) { set_error("bad right-hand side stride"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = pick_stream(d, stream);
    PoolBuf tmp;  // returned to the pool with an event: the call stays asynchronous
    RLG_TRY(tmp.alloc(nrhs * ch->n * 8, s));
    ch->use.add(s);
    return chol_solve<double>(s, ch->n, static_cast<const double *>(ch->a), ch->lda, nrhs, b, ldb, x,
                              ldx, tmp.as<double>());
}

int rlg_chol_size(const rlg_chol *ch, size_t *n) {
    if (!ch || !n) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    *n = ch->n;
    return RLG_OK;
}

int rlg_chol_unpack_f64(const rlg_chol *ch, double *l, size_t ld) { return chol_unpack<double>(ch, l, ld); }
int rlg_chol_unpack_f32(const rlg_chol *ch, float *l, size_t ld) { return chol_unpack<float>(ch, l, ld); }

int rlg_chol_det_f64(const rlg_chol *ch, double *det) {
    if (!ch || ch->elem != 8 || !det) { set_error("bad argument"); return RLG_E_INVALID_ARG; }
    const size_t n = ch->n;
    std::vector<double> diag(n);
    Device *d;
    RLG_TRY(current_device(&d));
    if (n)
        RLG_HIP(hipMemcpy2DAsync(diag.data(), 8, ch->a, (ch->lda + 1) * 8, 8, n,
                                 hipMemcpyDeviceToHost, d->stream));
    RLG_HIP(hipStreamSynchronize(d->stream));
    double l_det = 1.0;  // cholesky.rs:175-180
    for (size_t i = 0; i < n; ++i) l_det = l_det * diag[i];
    *det = l_det * l_det;
    return RLG_OK;
}

int rlg_chol_inverse_f64(const rlg_chol *ch, double *inv, size_t ld) {
    if (!ch || ch->elem != 8 || (!inv && ch->n)) { set_error("bad argument"); return RLG_E_INVALID_ARG; }
    const size_t n = ch->n;
    if (n == 0) return RLG_OK;
    std::vector<double> e(n * n, 0.0), x(n * n);
    for (size_t i = 0; i < n; ++i) e[i * n + i] = 1.0;
    RLG_TRY(rlg_chol_solve_f64(ch, n, n, e.data(), n, x.data(), n));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) inv[j * ld + i] = x[i * n + j];
    return RLG_OK;
}

void rlg_chol_free(rlg_chol *ch) {
    if (!ch) return;
    ch->own.release_after(ch->use.all());
    delete ch;
}

}  // extern "C"
