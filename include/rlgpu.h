/*
 * rlgpu: MI355X (gfx950) backend for rulinalg's dense GEMM and
 * factor-and-solve calls, exported as a plain C ABI.
 *
 * Conventions (the reference's, src/matrix/mod.rs:45-79):
 *   - matrices are row-major; element (i, j) of an operand with row stride
 *     `ld` is p[i * ld + j], ld >= cols (a MatrixSlice's row_stride);
 *   - a block of `nrhs` right-hand sides is `nrhs` vectors of length n, vector
 *     r at b + r * ldb (ldb >= n): one Vector<T> per reference solve() call;
 *   - entry points ending in _dev take device pointers and a hipStream_t
 *     passed as void* (NULL = the library's stream for the current device);
 *     all others take host pointers and are synchronous, like the reference.
 *   - status codes: RLG_OK and then the reference's ErrorKind in declaration
 *     order (src/error.rs:21-34); RLG_E_SHAPE stands for the reference's
 *     assert! panics, which cannot unwind across a C boundary.
 *   - rlg_last_error_message() gives a thread-local description of the last
 *     failure on the calling thread.
 *
 * Factor handles own device memory and are immutable after decompose, so
 * concurrent *_solve calls on one handle from several threads are allowed
 * (each call uses its own workspace), as `&self` allows in the reference.
 */
#ifndef RLGPU_H
#define RLGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rlg_status {
    RLG_OK = 0,
    RLG_E_INVALID_ARG = 1,         /* ErrorKind::InvalidArg */
    RLG_E_DECOMP_FAILURE = 2,      /* ErrorKind::DecompFailure */
    RLG_E_ALGEBRA_FAILURE = 3,     /* ErrorKind::AlgebraFailure */
    RLG_E_DIV_BY_ZERO = 4,         /* ErrorKind::DivByZero */
    RLG_E_SCALAR_CONVERSION = 5,   /* ErrorKind::ScalarConversionFailure */
    RLG_E_INVALID_PERMUTATION = 6, /* ErrorKind::InvalidPermutation */
    RLG_E_SHAPE = 7,               /* a reference assert!: dimension mismatch */
    RLG_E_DEVICE = 8,              /* HIP runtime or kernel failure */
    RLG_E_OOM = 9,                 /* device allocation failed / over budget */
    RLG_E_NOT_CONVERGED = 10       /* mixed solve stalled with fallback off */
} rlg_status;

/* ---------------------------------------------------------------- runtime */

const char *rlg_version(void);
const char *rlg_last_error_message(void);
/* Select the device used by later calls on this thread (default 0). */
int rlg_set_device(int device);
/* Block until all work the library queued on the current device is done. */
int rlg_synchronize(void);

/* Device memory helpers, so that callers without another GPU runtime can
 * keep operands resident.  Copies are synchronous. */
int rlg_dev_alloc(size_t bytes, void **ptr);
int rlg_dev_free(void *ptr);
int rlg_dev_upload(void *dst, const void *src, size_t bytes);
int rlg_dev_download(void *dst, const void *src, size_t bytes);
/* Device-to-device copy, asynchronous on `stream` (NULL: the library's). */
int rlg_dev_copy(void *dst, const void *src, size_t bytes, void *stream);

/* ---------------------------------------------------------------- GEMM
 * C = A * B for an (a_rows x a_cols) A and a (b_rows x b_cols) B: the
 * reference's `&a * &b` for Matrix and MatrixSlice operands
 * (src/matrix/mat_mul.rs:17-103, 137-143, 145-287).  a_cols != b_rows is the
 * reference's assert at mat_mul.rs:21 and returns RLG_E_SHAPE.  C is
 * a_rows x b_cols with row stride ldc (the reference returns it contiguous:
 * ldc = b_cols). */
int rlg_gemm_f64(size_t a_rows, size_t a_cols, const double *a, size_t lda,
                 size_t b_rows, size_t b_cols, const double *b, size_t ldb,
                 double *c, size_t ldc);
int rlg_gemm_f32(size_t a_rows, size_t a_cols, const float *a, size_t lda,
                 size_t b_rows, size_t b_cols, const float *b, size_t ldb,
                 float *c, size_t ldc);
/* Device form: C = alpha * A * B + beta * C, A m x k, B k x n. */
int rlg_gemm_f64_dev(size_t m, size_t k, size_t n, double alpha,
                     const double *a, size_t lda, const double *b, size_t ldb,
                     double beta, double *c, size_t ldc, void *stream);
int rlg_gemm_f32_dev(size_t m, size_t k, size_t n, float alpha,
                     const float *a, size_t lda, const float *b, size_t ldb,
                     float beta, float *c, size_t ldc, void *stream);
/* The same FP32 product on the FP16 matrix cores (no counterpart in the
 * reference; it stands behind the same `&a * &b` for f32, mat_mul.rs:57-67,
 * where the caller accepts FP32-level rather than FP32-exact rounding):
 * every row of A and column of B is scaled by a power of two and split as
 * hi + lo FP16, and C accumulates hi*hi + hi*lo + lo*hi in FP32.  Normwise
 * error of order 2^-22 ||A|| ||B|| plus FP32 accumulation.  Allocates a
 * packing workspace of about 8 (m + n) k bytes per call. */
int rlg_gemm_f32_split_dev(size_t m, size_t k, size_t n, float alpha,
                           const float *a, size_t lda, const float *b, size_t ldb,
                           float beta, float *c, size_t ldc, void *stream);

/* y = A * v for an m x n A: `&Matrix * &Vector` (src/matrix/impl_ops.rs:298-315).
 * v_len != n returns RLG_E_SHAPE (the assert at impl_ops.rs:304). */
int rlg_matvec_f64(size_t m, size_t n, const double *a, size_t lda,
                   size_t v_len, const double *v, double *y);

/* ---------------------------------------------------------------- permutations
 * PermutationMatrix * Matrix and Matrix * PermutationMatrix
 * (src/matrix/impl_permutation_mul.rs:53-127 and 129-204), with the
 * into-buffer semantics of permutation_matrix.rs:318-358: row i of `a`
 * becomes row perm[i] of `out` (rows form), column j of `a` becomes column
 * perm[j] of `out` (cols form).  perm holds the reference's Vec<usize>.
 * perm_len != rows (rows form) or != cols (cols form) returns RLG_E_SHAPE
 * (the asserts at permutation_matrix.rs:424-436); a perm that is not a
 * permutation of 0..perm_len returns RLG_E_INVALID_PERMUTATION
 * (validate_permutation, permutation_matrix.rs:438-467).  out must not
 * overlap a. */
int rlg_permute_rows_f64(size_t perm_len, const size_t *perm, size_t rows, size_t cols,
                         const double *a, size_t lda, double *out, size_t ldo);
int rlg_permute_rows_f32(size_t perm_len, const size_t *perm, size_t rows, size_t cols,
                         const float *a, size_t lda, float *out, size_t ldo);
int rlg_permute_cols_f64(size_t perm_len, const size_t *perm, size_t rows, size_t cols,
                         const double *a, size_t lda, double *out, size_t ldo);
int rlg_permute_cols_f32(size_t perm_len, const size_t *perm, size_t rows, size_t cols,
                         const float *a, size_t lda, float *out, size_t ldo);
/* Device forms (asynchronous on `stream`): perm is a device array of n
 * int64 that must be a permutation (not checked). */
int rlg_permute_rows_f64_dev(size_t n, const int64_t *perm, size_t cols, const double *a,
                             size_t lda, double *out, size_t ldo, void *stream);
int rlg_permute_cols_f64_dev(size_t rows, size_t n, const int64_t *perm, const double *a,
                             size_t lda, double *out, size_t ldo, void *stream);

/* ---------------------------------------------------------------- triangular
 * BaseMatrix::solve_u_triangular / solve_l_triangular
 * (src/matrix/base/mod.rs:1015-1024, 1058-1067) on nrhs right-hand sides
 * (vector r at b + r*ldb, solution at x + r*ldx; nrhs = 1 is the reference
 * call).  Only the upper (resp. lower) triangle is read.  b_len != cols or
 * rows != cols returns RLG_E_SHAPE (the asserts at base/mod.rs:1018 and
 * src/matrix/mod.rs:322); a diagonal entry with |m_ii| < epsilon returns
 * RLG_E_DIV_BY_ZERO (src/matrix/mod.rs:331-335, 375-378). */
int rlg_solve_u_triangular_f64(size_t rows, size_t cols, const double *u, size_t ldu,
                               size_t b_len, size_t nrhs, const double *b, size_t ldb,
                               double *x, size_t ldx);
int rlg_solve_l_triangular_f64(size_t rows, size_t cols, const double *l, size_t ldl,
                               size_t b_len, size_t nrhs, const double *b, size_t ldb,
                               double *x, size_t ldx);
int rlg_solve_u_triangular_f32(size_t rows, size_t cols, const float *u, size_t ldu,
                               size_t b_len, size_t nrhs, const float *b, size_t ldb,
                               float *x, size_t ldx);
int rlg_solve_l_triangular_f32(size_t rows, size_t cols, const float *l, size_t ldl,
                               size_t b_len, size_t nrhs, const float *b, size_t ldb,
                               float *x, size_t ldx);
/* Device form of both (upper != 0: U x = b, else L x = b); x may equal b.
 * Waits for the stream before returning, to report RLG_E_DIV_BY_ZERO. */
int rlg_solve_triangular_f64_dev(int upper, size_t n, const double *m, size_t ldm,
                                 size_t nrhs, const double *b, size_t ldb, double *x,
                                 size_t ldx, void *stream);

/* ---------------------------------------------------------------- LU
 * PartialPivLu (src/matrix/decomposition/lu.rs:130-300).  The factor lives
 * on the device: packed L (unit diagonal implied) and U, plus the pivots. */
typedef struct rlg_lu rlg_lu;

/* PartialPivLu::decompose (lu.rs:163-194).  rows != cols -> RLG_E_SHAPE
 * (lu.rs:165).  A pivot of magnitude below epsilon -> RLG_E_DIV_BY_ZERO
 * (lu.rs:179-183) and *out = NULL.  The caller's buffer is copied (the
 * reference consumes its matrix). */
int rlg_lu_decompose_f64(size_t rows, size_t cols, const double *a, size_t lda,
                         rlg_lu **out);
int rlg_lu_decompose_f32(size_t rows, size_t cols, const float *a, size_t lda,
                         rlg_lu **out);
/* Device form: factors a device matrix in place; the handle borrows `a`,
 * which must stay allocated until rlg_lu_free. */
int rlg_lu_decompose_f64_dev(size_t n, double *a, size_t lda, void *stream,
                             rlg_lu **out);

/* PartialPivLu::solve (lu.rs:231-244) for nrhs vectors at once.
 * b_len != n -> RLG_E_SHAPE (lu.rs:232).  x may alias b. */
int rlg_lu_solve_f64(const rlg_lu *lu, size_t b_len, size_t nrhs,
                     const double *b, size_t ldb, double *x, size_t ldx);
int rlg_lu_solve_f32(const rlg_lu *lu, size_t b_len, size_t nrhs,
                     const float *b, size_t ldb, float *x, size_t ldx);
int rlg_lu_solve_f64_dev(const rlg_lu *lu, size_t nrhs, const double *b,
                         size_t ldb, double *x, size_t ldx, void *stream);

int rlg_lu_size(const rlg_lu *lu, size_t *n);
/* The stored permutation `p` (lu.rs:190-193): the inverse of the product of
 * the row swaps, in PermutationMatrix's Vec<usize> convention
 * (src/matrix/permutation_matrix.rs:103-148, 412-416). */
int rlg_lu_perm(const rlg_lu *lu, size_t *perm);
/* The row chosen at each step (`curr_max_idx` at lu.rs:170-178), 0-based. */
int rlg_lu_ipiv(const rlg_lu *lu, int64_t *ipiv);
/* Packed factor, n x n with row stride ld (f64 handles / f32 handles). */
int rlg_lu_packed_f64(const rlg_lu *lu, double *out, size_t ld);
int rlg_lu_packed_f32(const rlg_lu *lu, float *out, size_t ld);
/* Decomposition::unpack (lu.rs:135-150): unit-lower L, upper U, and perm. */
int rlg_lu_unpack_f64(const rlg_lu *lu, double *l, double *u, size_t *perm);
/* PartialPivLu::det (lu.rs:291-300) and ::inverse (lu.rs:251-285). */
int rlg_lu_det_f64(const rlg_lu *lu, double *det);
int rlg_lu_inverse_f64(const rlg_lu *lu, double *inv, size_t ld);
void rlg_lu_free(rlg_lu *lu);

/* In-place FP64 LU factorization of a device matrix that allocates no device
 * memory.  No counterpart in the reference: it is for callers that cannot
 * call hipMalloc (code under HIP graph capture, serving loops and harnesses
 * that forbid allocation).  The factorization is rlg_lu_decompose_f64_dev's
 * (same kernels, same packed factor and pivots); its scratch is a static
 * device array of RLG_LU_NOALLOC_SCRATCH bytes that each GPU holds until
 * the process ends.  ipiv (n device int64) receives the step pivots as
 * rlg_lu_ipiv reports them; *info (device int64) receives n on success, the
 * first step whose pivot was below epsilon, or -1 if the panel exchange gave
 * up waiting.  Asynchronous on `stream` (NULL: the library's stream).  Every
 * call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * race.  A call whose scratch would exceed the array returns
 * RLG_E_INVALID_ARG. */
#define RLG_LU_NOALLOC_SCRATCH (1u << 20)
int rlg_lu_factor_f64_noalloc(size_t n, double *a, size_t lda, int64_t *ipiv,
                              int64_t *info, void *stream);

/* ---------------------------------------------------------------- FullPivLu
 * FullPivLu (src/matrix/decomposition/lu.rs:339-599): P A Q = L U with
 * complete pivoting.  The packed factor, p and q are identical to the
 * reference's (every element is computed with the reference's operations;
 * the pivot is the first maximal |value| in row-major order). */
typedef struct rlg_fplu rlg_fplu;

/* FullPivLu::decompose (lu.rs:383-432).  rows != cols -> RLG_E_SHAPE
 * (lu.rs:384-386).  A singular matrix is not an error: a zero pivot ends
 * the elimination (lu.rs:405-407).  Orders above 46340 -> RLG_E_INVALID_ARG. */
int rlg_fplu_decompose_f64(size_t rows, size_t cols, const double *a,
                           size_t lda, rlg_fplu **out);
int rlg_fplu_decompose_f32(size_t rows, size_t cols, const float *a,
                           size_t lda, rlg_fplu **out);
/* Factors the device matrix `a` in place; the handle refers to it. */
int rlg_fplu_decompose_f64_dev(size_t n, double *a, size_t lda, void *stream,
                               rlg_fplu **out);
/* FullPivLu::solve (lu.rs:458-464) on nrhs vectors.  b_len != n ->
 * RLG_E_SHAPE; a diagonal entry of U below epsilon in magnitude ->
 * RLG_E_DIV_BY_ZERO (back_substitution, src/matrix/mod.rs:330-336).  The
 * result is the reference's: the substitutions' result y is mapped through
 * `&q * y` (x[q[i]] = y[i]), which is the solution of A x = b when q is its
 * own inverse and otherwise a permutation of it. */
int rlg_fplu_solve_f64(const rlg_fplu *lu, size_t b_len, size_t nrhs,
                       const double *b, size_t ldb, double *x, size_t ldx);
int rlg_fplu_solve_f32(const rlg_fplu *lu, size_t b_len, size_t nrhs,
                       const float *b, size_t ldb, float *x, size_t ldx);
/* As rlg_fplu_solve_f64 but mapped through q^-1: the solution of A x = b
 * for every q (no counterpart in the reference, whose solve differs from it
 * when q is not its own inverse). */
int rlg_fplu_solve_true_f64(const rlg_fplu *lu, size_t b_len, size_t nrhs,
                            const double *b, size_t ldb, double *x,
                            size_t ldx);
/* Device form of both: true_solution = 0 gives rlg_fplu_solve_f64's result. */
int rlg_fplu_solve_f64_dev(const rlg_fplu *lu, size_t nrhs, const double *b,
                           size_t ldb, double *x, size_t ldx,
                           int true_solution, void *stream);
int rlg_fplu_size(const rlg_fplu *lu, size_t *n);
/* The stored p and q (lu.rs:425-429), n entries each (either may be NULL). */
int rlg_fplu_perms(const rlg_fplu *lu, size_t *p, size_t *q);
/* Elimination steps taken: n, or the step whose pivot was zero. */
int rlg_fplu_steps(const rlg_fplu *lu, size_t *steps);
/* Decomposition::unpack (lu.rs:348-361): unit-lower L and upper U, n x n
 * row-major (ld = n); p and q from rlg_fplu_perms. */
int rlg_fplu_unpack_f64(const rlg_fplu *lu, double *l, double *u);
int rlg_fplu_unpack_f32(const rlg_fplu *lu, float *l, float *u);
/* FullPivLu::rank (lu.rs:544-557), ::is_invertible (lu.rs:582-594),
 * ::det (lu.rs:505-517), ::inverse (lu.rs:472-497: RLG_E_DIV_BY_ZERO when
 * not invertible, else one solve per unit vector). */
int rlg_fplu_rank(const rlg_fplu *lu, size_t *rank);
int rlg_fplu_is_invertible(const rlg_fplu *lu, int *yes);
int rlg_fplu_det_f64(const rlg_fplu *lu, double *det);
int rlg_fplu_inverse_f64(const rlg_fplu *lu, double *inv, size_t ld);
void rlg_fplu_free(rlg_fplu *lu);

/* ---------------------------------------------------------------- Cholesky
 * Cholesky (src/matrix/decomposition/cholesky.rs:94-246).  Only the lower
 * triangle of the input is read (cholesky.rs:141-148). */
typedef struct rlg_chol rlg_chol;

/* Cholesky::decompose (cholesky.rs:116-169).  rows != cols -> RLG_E_SHAPE
 * (cholesky.rs:117); a diagonal below epsilon in magnitude or negative ->
 * RLG_E_DECOMP_FAILURE (cholesky.rs:151-158) and *out = NULL. */
int rlg_chol_decompose_f64(size_t rows, size_t cols, const double *a,
                           size_t lda, rlg_chol **out);
int rlg_chol_decompose_f32(size_t rows, size_t cols, const float *a,
                           size_t lda, rlg_chol **out);
int rlg_chol_decompose_f64_dev(size_t n, double *a, size_t lda, void *stream,
                               rlg_chol **out);
/* Cholesky::solve (cholesky.rs:194-202); b_len != n -> RLG_E_SHAPE. */
int rlg_chol_solve_f64(const rlg_chol *ch, size_t b_len, size_t nrhs,
                       const double *b, size_t ldb, double *x, size_t ldx);
int rlg_chol_solve_f32(const rlg_chol *ch, size_t b_len, size_t nrhs,
                       const float *b, size_t ldb, float *x, size_t ldx);
int rlg_chol_solve_f64_dev(const rlg_chol *ch, size_t nrhs, const double *b,
                           size_t ldb, double *x, size_t ldx, void *stream);
int rlg_chol_size(const rlg_chol *ch, size_t *n);
/* Decomposition::unpack (cholesky.rs:237-246): L with a zeroed upper part. */
int rlg_chol_unpack_f64(const rlg_chol *ch, double *l, size_t ld);
int rlg_chol_unpack_f32(const rlg_chol *ch, float *l, size_t ld);
/* Cholesky::det (cholesky.rs:175-180) and ::inverse (cholesky.rs:209-234). */
int rlg_chol_det_f64(const rlg_chol *ch, double *det);
int rlg_chol_inverse_f64(const rlg_chol *ch, double *inv, size_t ld);
void rlg_chol_free(rlg_chol *ch);

/* ---------------------------------------------------------------- mixed
 * Solve A X = B in FP64 by factoring a scaled copy of A in lower precision
 * and refining in FP64 (no counterpart in the reference, which is FP64 or
 * FP32 throughout; this stands behind the same solve() call).  The
 * residual R = B - A X, the update and the backward-error test run in FP64
 * on the device; when refinement stalls the solver falls back to the plain
 * FP64 factorization (rlg_lu_decompose_f64_dev + rlg_lu_solve_f64_dev). */
enum {
    RLG_PREC_F64 = 0,   /* plain FP64: no refinement */
    RLG_PREC_F32 = 1,   /* FP32 factor, FP32 MFMA update */
    RLG_PREC_F16 = 2,   /* FP32 panels, FP16 MFMA update, FP32 accumulate */
    RLG_PREC_BF16 = 3,  /* FP32 panels, BF16 MFMA update, FP32 accumulate */
    RLG_PREC_F16X3 = 4  /* FP32 panels; update operands split as hi + lo FP16
                           with power-of-two row scales, three FP16 MFMA
                           products (hi hi + hi lo + lo hi), FP32 accumulate:
                           an FP32-accurate update on the FP16 matrix cores */
};
enum {
    RLG_REFINE_IR = 0,     /* classical iterative refinement */
    RLG_REFINE_GMRES = 1   /* GMRES-IR: GMRES preconditioned by the factors */
};
enum {
    RLG_PIVOT_PARTIAL = 0,     /* partial pivoting, as PartialPivLu (lu.rs:170-178) */
    RLG_PIVOT_TOURNAMENT = 1,  /* tournament pivoting (TSLU) on the FP32 panels: its
                                  panels never wait for one another, so from order
                                  8192 on the trailing update overlaps them on a
                                  second stream */
    RLG_PIVOT_AUTO = 2         /* the default: tournament with FP32 or split-FP16
                                  updates, partial with FP16 or BF16 updates (whose
                                  factors, with tournament pivots, are too weak a
                                  preconditioner at kappa 1e2-1e4) */
};
typedef struct rlg_mixed_opts {
    int factor_prec;      /* RLG_PREC_* */
    int refine;           /* RLG_REFINE_* */
    double tol;           /* target normwise backward error (1e-13) */
    int max_iter;         /* refinement steps before declaring a stall */
    int fallback;         /* nonzero: fall back to plain FP64 on a stall */
    int pivot;            /* RLG_PIVOT_* for the low-precision factorization */
} rlg_mixed_opts;
typedef struct rlg_mixed_report {
    int iterations;        /* refinement steps taken (outer steps) */
    int inner_iterations;  /* GMRES iterations, summed */
    int fell_back;         /* 1 if the plain FP64 path produced X */
    int converged;         /* 1 if backward error <= tol on return */
    double backward_error; /* max over RHS of ||b-Ax||inf/(||A||inf||x||inf+||b||inf) */
    double factor_ms;      /* device time of the low-precision factorization */
    double refine_ms;      /* device time of the refinement loop */
    double fallback_ms;    /* device time spent in the FP64 fallback */
} rlg_mixed_report;
void rlg_mixed_opts_default(rlg_mixed_opts *o);
/* Device form: A (n x n, lda) and B are not modified; X receives the
 * solution.  Workspace is allocated per call. */
int rlg_solve_mixed_f64_dev(size_t n, const double *a, size_t lda, size_t nrhs,
                            const double *b, size_t ldb, double *x, size_t ldx,
                            const rlg_mixed_opts *opts, rlg_mixed_report *rep,
                            void *stream);
int rlg_solve_mixed_f64(size_t n, const double *a, size_t lda, size_t nrhs,
                        const double *b, size_t ldb, double *x, size_t ldx,
                        const rlg_mixed_opts *opts, rlg_mixed_report *rep);

/* Normwise backward error of a computed X on the device, FP64:
 * max_r ||b_r - A x_r||inf / (||A||inf ||x_r||inf + ||b_r||inf). */
int rlg_backward_error_f64_dev(size_t n, const double *a, size_t lda,
                               size_t nrhs, const double *b, size_t ldb,
                               const double *x, size_t ldx, double *eta,
                               void *stream);

/* ---------------------------------------------------------------- batched
 * Many independent systems of uneven order in one persistent launch.  Each
 * item is factored (LU with partial pivoting as PartialPivLu::decompose, or
 * Cholesky as Cholesky::decompose) and solved for its RHS, in place: `a` is
 * overwritten by the packed factor and `b` by the solution.  All pointers
 * are device pointers (row-major, ld = n, RHS vectors contiguous). */
enum { RLG_KIND_LU = 0, RLG_KIND_CHOL = 1 };
typedef struct rlg_batch_item {
    int kind;        /* RLG_KIND_* */
    int n;           /* order, 1..RLG_BATCH_MAX_N */
    int nrhs;        /* 1..RLG_BATCH_MAX_NRHS */
    int status;      /* out: RLG_OK / RLG_E_DIV_BY_ZERO / RLG_E_DECOMP_FAILURE */
    double *a;       /* n x n, overwritten by the factor */
    double *b;       /* nrhs x n, overwritten by the solution */
    int64_t *ipiv;   /* n entries (LU only, may be NULL for Cholesky) */
} rlg_batch_item;
#define RLG_BATCH_MAX_N 512
#define RLG_BATCH_MAX_NRHS 16
/* `items` is a device array of `count` items.  FP64 throughout. */
int rlg_batch_solve_f64_dev(size_t count, rlg_batch_item *items, void *stream);
/* Mixed form: each system is factored in FP32 from a copy of A and refined in
 * FP64 against the unmodified A; `a` is kept, `b` receives the solution.
 * Items whose refinement stalls are refactored in FP64 (status reports it
 * through *fallbacks). */
int rlg_batch_solve_mixed_dev(size_t count, rlg_batch_item *items,
                              int max_iter, double tol, size_t *fallbacks,
                              void *stream);

/* Autotune the batched solver: for each n-bucket (n <= 32, 64, 128, 256,
 * 512) time synthetic systems with each panel width (8, 16, 32) and keep the
 * fastest for later batched calls.  Writes one line per bucket into report
 * (may be NULL). */
int rlg_batch_autotune(int systems_per_bucket, char *report, size_t len);

/* The device part of rlg_batch_solve_f64_dev (work-counter reset and the
 * persistent kernel) for a fixed item array, captured once into a HIP graph
 * and replayed.  The data the items point to may change between replays;
 * the item array, orders and shapes may not.  Replays on one stream are
 * ordered; overlapping replays race on the shared work counter. */
typedef struct rlg_batch_graph rlg_batch_graph;
int rlg_batch_graph_create(size_t count, rlg_batch_item *items,
                           rlg_batch_graph **out);
int rlg_batch_graph_launch(rlg_batch_graph *g, void *stream);
void rlg_batch_graph_destroy(rlg_batch_graph *g);

/* ---------------------------------------------------------------- serving
 * Continuous-batching request loop.  Requests are admitted while device
 * memory under the budget and a stream slot are free; factors can be kept
 * resident under a handle for later right-hand sides. */
typedef struct rlg_server rlg_server;
typedef struct rlg_server_opts {
    size_t mem_budget_bytes;   /* device bytes the server may hold */
    int slots;                 /* concurrent in-flight requests (streams) */
    int mixed;                 /* nonzero: mixed-precision refined solves */
    double tol;                /* backward-error target for mixed solves */
} rlg_server_opts;
typedef struct rlg_request {
    int kind;            /* RLG_KIND_LU / RLG_KIND_CHOL */
    int n, nrhs;
    const double *a;     /* host, n x n row-major; NULL to reuse `factor_id` */
    const double *b;     /* host, nrhs x n */
    int64_t factor_id;   /* >0: keep (a != NULL) or reuse (a == NULL) a factor */
} rlg_request;
typedef struct rlg_result {
    int64_t ticket;
    int status;
    int fell_back;
    double backward_error;
    double latency_ms;   /* submit -> result ready */
    double *x;           /* host buffer owned by the server until released */
} rlg_result;
void rlg_server_opts_default(rlg_server_opts *o);
int rlg_server_create(const rlg_server_opts *o, rlg_server **out);
/* Non-blocking: queues the request and returns its ticket.  The host
 * buffers a and b are read when the request is admitted and must stay valid
 * until rlg_server_wait has returned for the ticket. */
int rlg_server_submit(rlg_server *s, const rlg_request *r, int64_t *ticket);
/* Blocks until the ticket's result is ready; `out->x` stays valid until
 * rlg_server_release(s, ticket). */
int rlg_server_wait(rlg_server *s, int64_t ticket, rlg_result *out);
int rlg_server_release(rlg_server *s, int64_t ticket);
int rlg_server_evict(rlg_server *s, int64_t factor_id);
void rlg_server_destroy(rlg_server *s);

/* ---------------------------------------------------------------- profiling
 * Per-kernel-class device time, measured with HIP events recorded around
 * every launch on the stream the kernel runs on.  Off by default. */
int rlg_prof_enable(int on);
int rlg_prof_reset(void);
/* Returns the number of launches, total milliseconds and total algorithmic
 * work (flops for GEMM classes: 2*m*n*k per launch; 0 where not counted) of
 * a kernel class ("gemm_f64", "gemm_f16", "lu_panel", "laswp", "trsm",
 * "trsv", ...).  Any output pointer may be NULL. */
int rlg_prof_get(const char *name, int64_t *launches, double *total_ms,
                 double *work);
/* Writes "name launches total_ms work\n" lines into buf (truncated to len). */
int rlg_prof_dump(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* RLGPU_H */
