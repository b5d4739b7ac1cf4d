/*
 * CPU oracle for the dense GEMM + factor-and-solve path of rulinalg 0.4.2.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by the
 * tests, by the package's smoke check and by the CPU-baseline leg of the
 * benchmark, never by the product library.  Nothing in rulinalg_amd/ calls it.
 *
 * Every function restates one routine of the reference line for line, with
 * the same operation order, so that with IEEE-754 binary64/binary32
 * round-to-nearest arithmetic and no FMA contraction (build with
 * -ffp-contract=off, no -ffast-math) the results are bit-identical to the Rust
 * code.  Rust performs no reassociation and no implicit FMA, and `sqrt` and
 * `/` are correctly rounded in both languages.
 *
 * Layout: row-major, element (i, j) at a[i * ld + j]
 * (src/matrix/impl_ops.rs:70-79, src/matrix/base/impl_base.rs:17-19).
 *
 * Status codes follow the order of ErrorKind (src/error.rs:21-34), plus one
 * code for the reference's shape panics.
 *
 * Pinning (what fixes each function beyond this restatement):
 *   dot, lu_decompose, lu_solve, forward/back substitutions, chol_decompose,
 *   chol_solve, transpose_back_substitution, permutation helpers, matvec,
 *   det/inverse: the reference's known-answer tests (tests/golden/kats.json),
 *   LAPACK dgetrf / dpotrf run on the same inputs through scipy.linalg
 *   (lu_factor, cholesky: tests/test_oracle.py::test_lu_matches_lapack,
 *   test_chol_matches_lapack),
 *   reconstruction and backward-error identities, and brute force on tiny
 *   cases (tests/test_oracle.py).
 *   matmul: the reference's exact small products (mat_mul.rs:293-427) and a
 *   compensated-sum product; the rounding order of matrixmultiply::dgemm
 *   itself is not available here, so bitwise GEMM parity is "parity
 *   unpinned" and GEMM parity is by tolerance.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

enum {
    ORC_OK = 0,
    ORC_E_INVALID_ARG = 1,      /* ErrorKind::InvalidArg */
    ORC_E_DECOMP_FAILURE = 2,   /* ErrorKind::DecompFailure */
    ORC_E_ALGEBRA_FAILURE = 3,  /* ErrorKind::AlgebraFailure */
    ORC_E_DIV_BY_ZERO = 4,      /* ErrorKind::DivByZero */
    ORC_E_SCALAR_CONVERSION = 5,
    ORC_E_INVALID_PERMUTATION = 6,
    ORC_E_SHAPE = 7             /* a reference assert!/panic */
};

/* ------------------------------------------------------------------------ */
/* Permutation helpers (type independent)                                    */
/* ------------------------------------------------------------------------ */

/* PermutationMatrix::inverse (src/matrix/permutation_matrix.rs:137-148):
 * inv[perm[source]] = source. */
void orc_perm_inverse(size_t n, const size_t *perm, size_t *inv)
{
    for (size_t source = 0; source < n; ++source)
        inv[perm[source]] = source;
}

/* PermutationMatrix::parity via permute_by_swap
 * (src/matrix/permutation_matrix.rs:216-234, 479-509).  Returns +1 for even,
 * -1 for odd.  Works on a copy, so `perm` is untouched. */
int orc_perm_parity(size_t n, const size_t *perm_in)
{
    size_t *perm = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
    memcpy(perm, perm_in, n * sizeof(size_t));
    int is_even = 1;
    for (size_t i = 0; i < n; ++i) {
        size_t target = perm[i];
        while (i != target) {
            size_t new_target = perm[target];
            is_even = !is_even;          /* the swap(i, target) callback */
            perm[target] = target;
            target = new_target;
        }
        perm[i] = i;
    }
    free(perm);
    return is_even ? 1 : -1;
}

/* ------------------------------------------------------------------------ */
/* Typed routines, instantiated for double (f64) and float (f32)             */
/* ------------------------------------------------------------------------ */

#define ORC_DEFINE(T, SUF, EPS, SQRT)                                                \
                                                                               \
/* utils::dot (src/utils.rs:20-51): eight running partial sums over chunks   \
 * of eight, combined as s=(s+p0)+p4, (s+p1)+p5, (s+p2)+p6, (s+p3)+p7, then  \
 * a sequential tail. */                                                       \
T orc_dot_##SUF(const T *u, const T *v, size_t len)                            \
{                                                                              \
    T s = 0, p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;   \
    size_t off = 0;                                                            \
    while (len - off >= 8) {                                                   \
        p0 = p0 + u[off + 0] * v[off + 0];                                     \
        p1 = p1 + u[off + 1] * v[off + 1];                                     \
        p2 = p2 + u[off + 2] * v[off + 2];                                     \
        p3 = p3 + u[off + 3] * v[off + 3];                                     \
        p4 = p4 + u[off + 4] * v[off + 4];                                     \
        p5 = p5 + u[off + 5] * v[off + 5];                                     \
        p6 = p6 + u[off + 6] * v[off + 6];                                     \
        p7 = p7 + u[off + 7] * v[off + 7];                                     \
        off += 8;                                                              \
    }                                                                          \
    s = s + p0 + p4;                                                           \
    s = s + p1 + p5;                                                           \
    s = s + p2 + p6;                                                           \
    s = s + p3 + p7;                                                           \
    for (; off < len; ++off)                                                   \
        s = s + u[off] * v[off];                                               \
    return s;                                                                  \
}                                                                              \
                                                                               \
/* Generic (non-BLAS) product of mat_mul_general! (src/matrix/mat_mul.rs:     \
 * 76-98): i-k-j loop, c[i][j] = c[i][j] + a[i][k]*b[k][j], C contiguous p x r.\
 * Strided operands follow MatrixSlice row_stride (src/matrix/slice.rs:15-33).\
 * Returns ORC_E_SHAPE for the `self.cols == m.rows` assert (line 21), which  \
 * a flat C signature cannot express, so the caller passes both. */           \
int orc_matmul_##SUF(size_t p, size_t q, size_t q2, size_t r,                  \
                     const T *a, size_t lda, const T *b, size_t ldb, T *c)     \
{                                                                              \
    if (q != q2) return ORC_E_SHAPE;                                           \
    for (size_t i = 0; i < p * r; ++i) c[i] = 0;                               \
    for (size_t i = 0; i < p; ++i)                                             \
        for (size_t k = 0; k < q; ++k)                                         \
            for (size_t j = 0; j < r; ++j)                                     \
                c[i * r + j] = c[i * r + j] + a[i * lda + k] * b[k * ldb + j]; \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* &Matrix * &Vector (src/matrix/impl_ops.rs:298-315): y_i = dot(row_i, v). */ \
void orc_matvec_##SUF(size_t rows, size_t cols, const T *a, const T *v, T *y)  \
{                                                                              \
    for (size_t i = 0; i < rows; ++i)                                          \
        y[i] = orc_dot_##SUF(a + i * cols, v, cols);                           \
}                                                                              \
                                                                               \
/* PartialPivLu::decompose (src/matrix/decomposition/lu.rs:163-194) with     \
 * gaussian_elimination (lu.rs:603-616) and BaseMatrixMut::swap_rows         \
 * (src/matrix/base/mod.rs:1394-1417).  `lu` (n x n, row-major, ld = n) is   \
 * overwritten with the packed factors.  Outputs:                            \
 *   perm  - the stored `p`, i.e. the inverse of the swap-built permutation  \
 *           (lu.rs:190-193);                                                \
 *   ipiv  - optional, the pivot row chosen at each step (`curr_max_idx`);   \
 *   margin- optional, per step |max| - |runner-up| over candidates, for     \
 *           near-tie diagnostics (0 on exact ties).                          \
 * On failure returns ORC_E_DIV_BY_ZERO and *fail_step = the failing index,  \
 * leaving `lu` as the reference leaves its (dropped) working copy. */       \
int orc_lu_decompose_##SUF(size_t n, T *lu, size_t *perm, int64_t *ipiv,       \
                           T *margin, size_t *fail_step)                       \
{                                                                              \
    size_t *raw = (size_t *)malloc((n ? n : 1) * sizeof(size_t));              \
    for (size_t i = 0; i < n; ++i) raw[i] = i;                                 \
    for (size_t index = 0; index < n; ++index) {                               \
        size_t curr_max_idx = index;                                           \
        T curr_max = lu[curr_max_idx * n + curr_max_idx];                      \
        T second = (T)-1;                                                      \
        for (size_t i = curr_max_idx + 1; i < n; ++i) {                        \
            T v = lu[i * n + index];                                           \
            if (fabs((double)v) > fabs((double)curr_max)) {                    \
                if (margin) second = (T)fabs((double)curr_max);                \
                curr_max = v;                                                  \
                curr_max_idx = i;                                              \
            } else if (margin && (T)fabs((double)v) > second) {                \
                second = (T)fabs((double)v);                                   \
            }                                                                  \
        }                                                                      \
        if (margin)                                                            \
            margin[index] = second < 0 ? (T)INFINITY                           \
                                       : (T)fabs((double)curr_max) - second;   \
        if (ipiv) ipiv[index] = (int64_t)curr_max_idx;                         \
        if ((T)fabs((double)curr_max) < EPS) {                                 \
            if (fail_step) *fail_step = index;                                 \
            free(raw);                                                         \
            return ORC_E_DIV_BY_ZERO;                                          \
        }                                                                      \
        if (index != curr_max_idx) {                                           \
            for (size_t j = 0; j < n; ++j) {                                   \
                T t = lu[index * n + j];                                       \
                lu[index * n + j] = lu[curr_max_idx * n + j];                  \
                lu[curr_max_idx * n + j] = t;                                  \
            }                                                                  \
            size_t t = raw[index];                                             \
            raw[index] = raw[curr_max_idx];                                    \
            raw[curr_max_idx] = t;                                             \
        }                                                                      \
        /* gaussian_elimination(&mut lu, index) */                            \
        T piv_val = lu[index * n + index];                                     \
        for (size_t i = index + 1; i < n; ++i) {                               \
            T mult = lu[i * n + index] / piv_val;                              \
            lu[i * n + index] = mult;                                          \
            for (size_t j = index + 1; j < n; ++j)                             \
                lu[i * n + j] = lu[i * n + j] - mult * lu[index * n + j];      \
        }                                                                      \
    }                                                                          \
    orc_perm_inverse(n, raw, perm);                                            \
    free(raw);                                                                 \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* lu_forward_substitution (lu.rs:624-642): x[i] -= fold_{j<i}(l[i][j]*x[j]) \
 * with a sequential left fold from zero; unit diagonal implied. */          \
void orc_lu_forward_substitution_##SUF(size_t n, const T *lu, T *x)            \
{                                                                              \
    for (size_t i = 1; i < n; ++i) {                                           \
        T adjustment = 0;                                                      \
        for (size_t j = 0; j < i; ++j)                                         \
            adjustment = adjustment + lu[i * n + j] * x[j];                    \
        x[i] = x[i] - adjustment;                                              \
    }                                                                          \
}                                                                              \
                                                                               \
/* back_substitution (src/matrix/mod.rs:318-357). */                          \
int orc_back_substitution_##SUF(size_t n, const T *u, size_t ldu, T *x)        \
{                                                                              \
    for (size_t ii = n; ii-- > 0;) {                                           \
        T divisor = u[ii * ldu + ii];                                          \
        if ((T)fabs((double)divisor) < EPS) return ORC_E_DIV_BY_ZERO;          \
        T d = orc_dot_##SUF(u + ii * ldu + ii + 1, x + ii + 1, n - (ii + 1));  \
        x[ii] = (x[ii] - d) / divisor;                                         \
    }                                                                          \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* forward_substitution (src/matrix/mod.rs:363-398). */                       \
int orc_forward_substitution_##SUF(size_t n, const T *l, size_t ldl, T *x)     \
{                                                                              \
    for (size_t i = 0; i < n; ++i) {                                           \
        T divisor = l[i * ldl + i];                                            \
        if ((T)fabs((double)divisor) < EPS) return ORC_E_DIV_BY_ZERO;          \
        T d = orc_dot_##SUF(l + i * ldl, x, i);                                \
        x[i] = (x[i] - d) / divisor;                                           \
    }                                                                          \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* PartialPivLu::solve (lu.rs:231-244): y = &p * b, which is               \
 * permute_vector_into_buffer (permutation_matrix.rs:369-382,               \
 * impl_permutation_mul.rs:21-41): y[perm[i]] = b[i]; then the two          \
 * substitutions.  `x` receives the solution; b is not modified. */          \
int orc_lu_solve_##SUF(size_t n, const T *lu, const size_t *perm,             \
                       const T *b, T *x)                                       \
{                                                                              \
    for (size_t i = 0; i < n; ++i) x[perm[i]] = b[i];                          \
    orc_lu_forward_substitution_##SUF(n, lu, x);                               \
    return orc_back_substitution_##SUF(n, lu, n, x);                           \
}                                                                              \
                                                                               \
/* PartialPivLu::det (lu.rs:291-300): product of the diagonal folded from  \
 * one, times the parity of p. */                                           \
T orc_lu_det_##SUF(size_t n, const T *lu, const size_t *perm)                  \
{                                                                              \
    T u_det = 1;                                                               \
    for (size_t i = 0; i < n; ++i) u_det = u_det * lu[i * n + i];              \
    T p_det = orc_perm_parity(n, perm) > 0 ? (T)1 : (T)0 - (T)1;               \
    return p_det * u_det;                                                      \
}                                                                              \
                                                                               \
/* PartialPivLu::inverse (lu.rs:251-285): solve against e_i, column i.     \
 * inv is n x n row-major. */                                               \
int orc_lu_inverse_##SUF(size_t n, const T *lu, const size_t *perm, T *inv)    \
{                                                                              \
    T *e = (T *)calloc(n ? n : 1, sizeof(T));                                  \
    T *col = (T *)malloc((n ? n : 1) * sizeof(T));                             \
    int st = ORC_OK;                                                           \
    for (size_t i = 0; i < n && st == ORC_OK; ++i) {                           \
        e[i] = 1;                                                              \
        st = orc_lu_solve_##SUF(n, lu, perm, e, col);                          \
        for (size_t j = 0; j < n; ++j) inv[j * n + i] = col[j];                \
        e[i] = 0;                                                              \
    }                                                                          \
    free(e); free(col);                                                        \
    return st;                                                                 \
}                                                                              \
                                                                               \
/* Cholesky::decompose (src/matrix/decomposition/cholesky.rs:116-169):      \
 * left-looking gaxpy Cholesky reading the lower triangle only.  On success \
 * `a` holds L in its lower triangle and the untouched input above it, as   \
 * the reference does; *fail_col receives the failing column on error. */   \
int orc_chol_decompose_##SUF(size_t n, T *a, size_t *fail_col)                 \
{                                                                              \
    for (size_t j = 0; j < n; ++j) {                                           \
        if (j > 0) {                                                           \
            for (size_t k = j; k < n; ++k) {                                   \
                T kj_dot = orc_dot_##SUF(a + k * n, a + j * n, j);             \
                a[k * n + j] = a[k * n + j] - kj_dot;                          \
            }                                                                  \
        }                                                                      \
        T diagonal = a[j * n + j];                                             \
        if ((T)fabs((double)diagonal) < EPS) {                                 \
            if (fail_col) *fail_col = j;                                       \
            return ORC_E_DECOMP_FAILURE;                                       \
        } else if (diagonal < (T)0) {                                          \
            if (fail_col) *fail_col = j;                                       \
            return ORC_E_DECOMP_FAILURE;                                       \
        }                                                                      \
        T divisor = SQRT(diagonal);                          \
        for (size_t k = j; k < n; ++k)                                         \
            a[k * n + j] = a[k * n + j] / divisor;                             \
    }                                                                          \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* transpose_back_substitution (cholesky.rs:329-361): column sweep,        \
 * x[i] /= l_ii, then x[j] = x[j] - x[i]*row[j] for j < i. */                \
int orc_transpose_back_substitution_##SUF(size_t n, const T *l, size_t ldl,    \
                                          T *x)                                \
{                                                                              \
    for (size_t i = n; i-- > 0;) {                                             \
        const T *row = l + i * ldl;                                            \
        T diagonal = l[i * ldl + i];                                           \
        if ((T)fabs((double)diagonal) < EPS) return ORC_E_DIV_BY_ZERO;         \
        x[i] = x[i] / diagonal;                                                \
        for (size_t j = 0; j < i; ++j) x[j] = x[j] - x[i] * row[j];            \
    }                                                                          \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* Cholesky::solve (cholesky.rs:194-202): forward_substitution then         \
 * transpose_back_substitution.  x receives the solution. */                 \
int orc_chol_solve_##SUF(size_t n, const T *l, const T *b, T *x)               \
{                                                                              \
    for (size_t i = 0; i < n; ++i) x[i] = b[i];                                \
    int st = orc_forward_substitution_##SUF(n, l, n, x);                       \
    if (st != ORC_OK) return st;                                               \
    return orc_transpose_back_substitution_##SUF(n, l, n, x);                  \
}                                                                              \
                                                                               \
/* Cholesky::det (cholesky.rs:175-180). */                                    \
T orc_chol_det_##SUF(size_t n, const T *l)                                     \
{                                                                              \
    T l_det = 1;                                                               \
    for (size_t i = 0; i < n; ++i) l_det = l_det * l[i * n + i];               \
    return l_det * l_det;                                                      \
}                                                                              \
                                                                               \
/* Cholesky::inverse (cholesky.rs:209-234). */                                \
int orc_chol_inverse_##SUF(size_t n, const T *l, T *inv)                       \
{                                                                              \
    T *e = (T *)calloc(n ? n : 1, sizeof(T));                                  \
    T *col = (T *)malloc((n ? n : 1) * sizeof(T));                             \
    int st = ORC_OK;                                                           \
    for (size_t i = 0; i < n && st == ORC_OK; ++i) {                           \
        e[i] = 1;                                                              \
        st = orc_chol_solve_##SUF(n, l, e, col);                               \
        for (size_t j = 0; j < n; ++j) inv[j * n + i] = col[j];                \
        e[i] = 0;                                                              \
    }                                                                          \
    free(e); free(col);                                                        \
    return st;                                                                 \
}                                                                              \
                                                                               \
/* PartialPivLu::unpack (lu.rs:135-150) with unit_lower_triangular_part    \
 * (lu.rs:644-663) and nullify_lower_triangular_part                         \
 * (src/internal_utils.rs:5-12). */                                          \
void orc_lu_unpack_##SUF(size_t n, const T *lu, T *l, T *u)                    \
{                                                                              \
    for (size_t i = 0; i < n; ++i)                                             \
        for (size_t j = 0; j < n; ++j) {                                       \
            T v = lu[i * n + j];                                               \
            l[i * n + j] = j < i ? v : (j == i ? (T)1 : (T)0);                 \
            u[i * n + j] = j < i ? (T)0 : v;                                   \
        }                                                                      \
}                                                                              \
                                                                               \
/* Cholesky::unpack (cholesky.rs:237-246) with                               \
 * nullify_upper_triangular_part (src/internal_utils.rs:14-21). */           \
void orc_chol_unpack_##SUF(size_t n, T *l)                                     \
{                                                                              \
    for (size_t i = 0; i < n; ++i)                                             \
        for (size_t j = i + 1; j < n; ++j) l[i * n + j] = 0;                   \
}

ORC_DEFINE(double, f64, DBL_EPSILON, sqrt)
ORC_DEFINE(float, f32, FLT_EPSILON, sqrtf)

/* ------------------------------------------------------------------------ */
/* FullPivLu (src/matrix/decomposition/lu.rs:339-599)                        */
/* ------------------------------------------------------------------------ */

#define ORC_DEFINE_FPLU(T, SUF, EPS)                                            \
                                                                               \
/* FullPivLu::decompose (lu.rs:383-432) with select_pivot (lu.rs:364-380),   \
 * gaussian_elimination (lu.rs:603-616), BaseMatrixMut::swap_rows and        \
 * swap_cols (src/matrix/base/mod.rs:1394-1462).  select_pivot scans the     \
 * bottom-right corner row by row with a strict `>` starting from           \
 * (index, index), so the pivot is the first maximal |value| in row-major    \
 * order.  A zero pivot ends the loop (lu.rs:405-407).  Outputs: `lu` the    \
 * packed factors; p and q the stored permutations (the inverses of the     \
 * swap-built ones, lu.rs:425-429); prow/pcol (optional) the pivot row and   \
 * column of each step taken; *steps the number of steps taken. */           \
int orc_fplu_decompose_##SUF(size_t n, T *lu, size_t *p, size_t *q,            \
                             int64_t *prow, int64_t *pcol, size_t *steps)      \
{                                                                              \
    size_t *rp = (size_t *)malloc((n ? n : 1) * sizeof(size_t));               \
    size_t *rq = (size_t *)malloc((n ? n : 1) * sizeof(size_t));               \
    for (size_t i = 0; i < n; ++i) { rp[i] = i; rq[i] = i; }                   \
    size_t index = 0;                                                          \
    for (; index < n; ++index) {                                               \
        size_t piv_row = index, piv_col = index;                               \
        T piv_val = lu[index * n + index];                                     \
        for (size_t row = index; row < n; ++row)                               \
            for (size_t col = index; col < n; ++col) {                         \
                T val = lu[row * n + col];                                     \
                if ((T)fabs((double)val) > (T)fabs((double)piv_val)) {         \
                    piv_val = val; piv_row = row; piv_col = col;               \
                }                                                              \
            }                                                                  \
        if ((T)fabs((double)piv_val) == (T)0) break;                           \
        if (prow) prow[index] = (int64_t)piv_row;                              \
        if (pcol) pcol[index] = (int64_t)piv_col;                              \
        if (index != piv_row)                                                  \
            for (size_t j = 0; j < n; ++j) {                                   \
                T t = lu[index * n + j];                                       \
                lu[index * n + j] = lu[piv_row * n + j];                       \
                lu[piv_row * n + j] = t;                                       \
            }                                                                  \
        if (index != piv_col)                                                  \
            for (size_t i = 0; i < n; ++i) {                                   \
                T t = lu[i * n + index];                                       \
                lu[i * n + index] = lu[i * n + piv_col];                       \
                lu[i * n + piv_col] = t;                                       \
            }                                                                  \
        { size_t t = rp[index]; rp[index] = rp[piv_row]; rp[piv_row] = t; }    \
        { size_t t = rq[index]; rq[index] = rq[piv_col]; rq[piv_col] = t; }    \
        T pv = lu[index * n + index];                                          \
        for (size_t i = index + 1; i < n; ++i) {                               \
            T mult = lu[i * n + index] / pv;                                   \
            lu[i * n + index] = mult;                                          \
            for (size_t j = index + 1; j < n; ++j)                             \
                lu[i * n + j] = lu[i * n + j] - mult * lu[index * n + j];      \
        }                                                                      \
    }                                                                          \
    if (steps) *steps = index;                                                 \
    orc_perm_inverse(n, rp, p);                                                \
    orc_perm_inverse(n, rq, q);                                                \
    free(rp); free(rq);                                                        \
    return ORC_OK;                                                             \
}                                                                              \
                                                                               \
/* FullPivLu::solve (lu.rs:458-464): forward substitution on &p * b, back   \
 * substitution, then &q * x (y[q[i]] = x[i]). */                            \
int orc_fplu_solve_##SUF(size_t n, const T *lu, const size_t *p,               \
                         const size_t *q, const T *b, T *x)                    \
{                                                                              \
    T *y = (T *)malloc((n ? n : 1) * sizeof(T));                               \
    for (size_t i = 0; i < n; ++i) y[p[i]] = b[i];                             \
    orc_lu_forward_substitution_##SUF(n, lu, y);                               \
    int st = orc_back_substitution_##SUF(n, lu, n, y);                         \
    if (st == ORC_OK)                                                          \
        for (size_t i = 0; i < n; ++i) x[q[i]] = y[i];                         \
    free(y);                                                                   \
    return st;                                                                 \
}                                                                              \
                                                                               \
/* FullPivLu::epsilon (lu.rs:596-598): |lu[0][0]| * eps, or eps for n = 0. */  \
static T orc_fplu_eps_##SUF(size_t n, const T *lu)                             \
{                                                                              \
    return (n ? (T)fabs((double)lu[0]) : (T)1) * EPS;                          \
}                                                                              \
                                                                               \
/* FullPivLu::rank (lu.rs:544-557): leading diagonal entries above eps. */   \
size_t orc_fplu_rank_##SUF(size_t n, const T *lu)                              \
{                                                                              \
    T eps = orc_fplu_eps_##SUF(n, lu);                                         \
    size_t rank = 0;                                                           \
    for (size_t i = 0; i < n; ++i) {                                           \
        if ((T)fabs((double)lu[i * n + i]) > eps) rank = rank + 1;             \
        else break;                                                            \
    }                                                                          \
    return rank;                                                               \
}                                                                              \
                                                                               \
/* FullPivLu::is_invertible (lu.rs:582-594): the last diagonal entry above   \
 * eps; true for n = 0. */                                                   \
int orc_fplu_is_invertible_##SUF(size_t n, const T *lu)                        \
{                                                                              \
    if (n == 0) return 1;                                                      \
    return (T)fabs((double)lu[(n - 1) * n + (n - 1)]) > orc_fplu_eps_##SUF(n, lu); \
}                                                                              \
                                                                               \
/* FullPivLu::det (lu.rs:505-517): diagonal product folded from one, times  \
 * the parities of p and q. */                                               \
T orc_fplu_det_##SUF(size_t n, const T *lu, const size_t *p, const size_t *q)  \
{                                                                              \
    T u_det = 1;                                                               \
    for (size_t i = 0; i < n; ++i) u_det = u_det * lu[i * n + i];              \
    T p_det = orc_perm_parity(n, p) > 0 ? (T)1 : (T)0 - (T)1;                  \
    T q_det = orc_perm_parity(n, q) > 0 ? (T)1 : (T)0 - (T)1;                  \
    return p_det * u_det * q_det;                                              \
}                                                                              \
                                                                               \
/* FullPivLu::inverse (lu.rs:472-497): DivByZero when not invertible, else   \
 * one solve per unit vector, stored as column i. */                         \
int orc_fplu_inverse_##SUF(size_t n, const T *lu, const size_t *p,             \
                           const size_t *q, T *inv)                            \
{                                                                              \
    if (!orc_fplu_is_invertible_##SUF(n, lu)) return ORC_E_DIV_BY_ZERO;        \
    T *e = (T *)calloc(n ? n : 1, sizeof(T));                                  \
    T *col = (T *)malloc((n ? n : 1) * sizeof(T));                             \
    int st = ORC_OK;                                                           \
    for (size_t i = 0; i < n && st == ORC_OK; ++i) {                           \
        e[i] = 1;                                                              \
        st = orc_fplu_solve_##SUF(n, lu, p, q, e, col);                        \
        for (size_t j = 0; j < n; ++j) inv[j * n + i] = col[j];                \
        e[i] = 0;                                                              \
    }                                                                          \
    free(e); free(col);                                                        \
    return st;                                                                 \
}

ORC_DEFINE_FPLU(double, f64, DBL_EPSILON)
ORC_DEFINE_FPLU(float, f32, FLT_EPSILON)

/* ------------------------------------------------------------------------ */
/* Baseline GEMM stand-in                                                    */
/* ------------------------------------------------------------------------ */

/* Packed, cache-blocked single-thread DGEMM used only as the CPU timing
 * baseline for `&a * &b` with f64 (the reference calls the external
 * matrixmultiply::dgemm, mat_mul.rs:57-67, whose source is not available).
 * C = A*B, row-major, C contiguous m x n.  Blocking: MC x KC panels of A and
 * KC x NC panels of B packed contiguously, 4x4 register microkernel. */
void orc_gemm_blocked_f64(size_t m, size_t k, size_t n, const double *a,
                          size_t lda, const double *b, size_t ldb, double *c)
{
    enum { MC = 64, KC = 256, NC = 512 };
    double *ap = (double *)aligned_alloc(64, sizeof(double) * MC * KC);
    double *bp = (double *)aligned_alloc(64, sizeof(double) * KC * NC);
    for (size_t i = 0; i < m * n; ++i) c[i] = 0.0;
    for (size_t jc = 0; jc < n; jc += NC) {
        size_t nc = n - jc < NC ? n - jc : NC;
        for (size_t pc = 0; pc < k; pc += KC) {
            size_t kc = k - pc < KC ? k - pc : KC;
            for (size_t p = 0; p < kc; ++p)
                for (size_t j = 0; j < nc; ++j)
                    bp[p * NC + j] = b[(pc + p) * ldb + jc + j];
            for (size_t ic = 0; ic < m; ic += MC) {
                size_t mc = m - ic < MC ? m - ic : MC;
                for (size_t i = 0; i < mc; ++i)
                    for (size_t p = 0; p < kc; ++p)
                        ap[i * KC + p] = a[(ic + i) * lda + pc + p];
                for (size_t i = 0; i < mc; i += 4) {
                    size_t mi = mc - i < 4 ? mc - i : 4;
                    for (size_t j = 0; j < nc; j += 4) {
                        size_t nj = nc - j < 4 ? nc - j : 4;
                        double acc[4][4] = {{0}};
                        for (size_t p = 0; p < kc; ++p) {
                            const double *bb = bp + p * NC + j;
                            for (size_t ii = 0; ii < mi; ++ii) {
                                double av = ap[(i + ii) * KC + p];
                                for (size_t jj = 0; jj < nj; ++jj)
                                    acc[ii][jj] += av * bb[jj];
                            }
                        }
                        for (size_t ii = 0; ii < mi; ++ii)
                            for (size_t jj = 0; jj < nj; ++jj)
                                c[(ic + i + ii) * n + jc + j + jj] += acc[ii][jj];
                    }
                }
            }
        }
    }
    free(ap);
    free(bp);
}
