// Internal launchers shared between the translation units of librlgpu.
// All matrices are row-major with explicit row strides; all launches are
// asynchronous on the given stream.
#pragma once

#include "common.hpp"

namespace rlg {

// C = alpha * op(A) * op(B) + beta * C on the MFMA pipes.
//   A is m x k; B is k x n, or n x k when trans_b (then op(B) = B^T).
//   lower_only: C is square (m == n) and only tiles touching the lower
//   triangle are computed (the SYRK shape of the Cholesky update).
// beta == 0 never reads C.
int gemm_f64(hipStream_t s, size_t m, size_t n, size_t k, double alpha,
             const double *a, size_t lda, const double *b, size_t ldb,
             double beta, double *c, size_t ldc, bool trans_b = false,
             bool lower_only = false);
int gemm_f32(hipStream_t s, size_t m, size_t n, size_t k, float alpha,
             const float *a, size_t lda, const float *b, size_t ldb,
             float beta, float *c, size_t ldc, bool trans_b = false,
             bool lower_only = false);
// FP32 operands rounded to FP16 (prec = RLG_PREC_F16) or BF16
// (RLG_PREC_BF16) on their way into LDS, products accumulated in FP32.
// With `pack` (at least gemm_lowp_pack_bytes(m, n, k) bytes of device
// memory) both operands are first rounded into k-contiguous packed copies
// and a plain NT kernel runs on them; without it the rounding happens on the
// way into LDS.
int gemm_lowp(hipStream_t s, int prec, size_t m, size_t n, size_t k,
              float alpha, const float *a, size_t lda, const float *b,
              size_t ldb, float beta, float *c, size_t ldc,
              bool trans_b = false, bool lower_only = false,
              void *pack = nullptr, size_t pack_bytes = 0, size_t lds_pad = 0);
size_t gemm_lowp_pack_bytes(size_t m, size_t n, size_t k);

template <class T>
inline int gemm_t(hipStream_t s, size_t m, size_t n, size_t k, T alpha,
                  const T *a, size_t lda, const T *b, size_t ldb, T beta, T *c,
                  size_t ldc, bool trans_b = false, bool lower_only = false);
template <>
inline int gemm_t<double>(hipStream_t s, size_t m, size_t n, size_t k,
                          double alpha, const double *a, size_t lda,
                          const double *b, size_t ldb, double beta, double *c,
                          size_t ldc, bool tb, bool lo) {
    return gemm_f64(s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tb, lo);
}
template <>
inline int gemm_t<float>(hipStream_t s, size_t m, size_t n, size_t k,
                         float alpha, const float *a, size_t lda,
                         const float *b, size_t ldb, float beta, float *c,
                         size_t ldc, bool tb, bool lo) {
    return gemm_f32(s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tb, lo);
}

// As gemm_t, with lds_pad bytes of dynamic LDS per workgroup that the kernel
// does not use: fewer of its workgroups then share a CU, which leaves room
// beside it for another kernel's workgroups (the general kernel only).
int gemm_f64_padded(hipStream_t s, size_t m, size_t n, size_t k, double alpha, const double *a,
                    size_t lda, const double *b, size_t ldb, double beta, double *c, size_t ldc,
                    bool trans_b, bool lower_only, size_t lds_pad);
int gemm_f32_padded(hipStream_t s, size_t m, size_t n, size_t k, float alpha, const float *a,
                    size_t lda, const float *b, size_t ldb, float beta, float *c, size_t ldc,
                    bool trans_b, bool lower_only, size_t lds_pad);
template <class T>
inline int gemm_t_padded(hipStream_t s, size_t m, size_t n, size_t k, T alpha, const T *a, size_t lda,
                         const T *b, size_t ldb, T beta, T *c, size_t ldc, bool tb, bool lo,
                         size_t lds_pad) {
    if constexpr (sizeof(T) == 8)
        return gemm_f64_padded(s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tb, lo, lds_pad);
    else
        return gemm_f32_padded(s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tb, lo, lds_pad);
}

// y = A x (FP64, one row per wave), A m x n.
int gemv_f64(hipStream_t s, size_t m, size_t n, const double *a, size_t lda,
             const double *x, double *y);

// ------------------------------------------------------------------ LU
// Right-looking blocked LU with partial pivoting of the n x n matrix `a`
// in place.  ipiv[k] receives the (0-based, absolute) row swapped with row
// k at step k.  *info (device) receives the first step whose pivot was
// below epsilon in magnitude, or n if none.  Also uses `work`, a device
// buffer of lu_work_bytes(n) bytes.
size_t lu_work_bytes(size_t n, size_t elem);
template <class T>
int lu_factor(hipStream_t s, size_t n, T *a, size_t lda, int64_t *ipiv,
              int64_t *info, void *work);
// Mixed-precision variant of the FP32 factorization whose trailing updates
// run on FP16/BF16 MFMA (prec = RLG_PREC_F16 / RLG_PREC_BF16).
// Its workspace is lu_lowp_work_bytes(n) bytes (lu_work_bytes plus room for
// the packed low-precision operands of the trailing update).
size_t lu_lowp_work_bytes(size_t n);
// `pivot`: RLG_PIVOT_PARTIAL (the FP32 form of k_lu_panel) or
// RLG_PIVOT_TOURNAMENT (tslu.hip); `prec` RLG_PREC_F32 keeps the trailing
// updates in FP32.
int lu_factor_lowp(hipStream_t s, int prec, int pivot, size_t n, float *a, size_t lda,
                   int64_t *ipiv, int64_t *info, void *work);

// Tournament-pivoting FP32 panel (tslu.hip), used by lu_factor_lowp: the
// workspace of tslu_work_bytes(n) bytes (for panels of up to n rows) is
// cleared by tslu_reset once per factorization.  tslu_panel factors the
// panel A[k:k+m, k:k+w] in place, writes the swaps ipiv[k..k+w) (the
// reference's sequential convention), the first tiny pivot into *info, and
// L11^-1 of the pivot block (64 x 64, zero outside w x w) into linv.
size_t tslu_work_bytes(size_t n);
size_t tslu_max_rows(size_t n);
int tslu_reset(hipStream_t s, void *work, size_t n);
int tslu_panel(hipStream_t s, float *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
               int64_t *info, size_t n, void *work, float *linv);

// Device flag raised by a panel barrier that gave up waiting (see lu.hip);
// nonzero after lu_factor means the factorization is invalid.
int *lu_timeout_flag(void *work, size_t n, size_t elem);

// Row-permutation map from the step pivots (host): raw[k] = the original
// row now at position k, i.e. the product of the swaps before the
// reference's final inverse (lu.rs:186-193).
void perm_from_ipiv(size_t n, const int64_t *ipiv, int64_t *raw);

// X = U^-1 L^-1 P B for nrhs vectors (vector r at b + r*ldb); x may alias b.
// `tmp` holds n * nrhs elements.
template <class T>
int lu_solve(hipStream_t s, size_t n, const T *lu, size_t lda,
             const int64_t *raw, size_t nrhs, const T *b, size_t ldb, T *x,
             size_t ldx, T *tmp);

// ------------------------------------------------------------------ FullPivLu
// LU with complete pivoting of the n x n matrix `a` in place (fplu.hip), in
// a workspace of fplu_work_bytes(n, elem) bytes; fplu_result then copies the
// step pivots (row and column of each step) and the number of steps taken
// to the host.
size_t fplu_work_bytes(size_t n, size_t elem);
template <class T>
int fplu_factor(hipStream_t s, size_t n, T *a, size_t lda, void *work);
int fplu_result(hipStream_t s, size_t n, void *work, size_t elem, int64_t *prow,
                int64_t *pcol, int64_t *steps);
// x[r][perm[i]] = y[r][i] for nrhs vectors of length n.
template <class T>
int scatter_vec(hipStream_t s, size_t n, size_t nrhs, const int64_t *perm, const T *y,
                size_t ldy, T *x, size_t ldx);

// ------------------------------------------------------------------ Cholesky
size_t chol_work_bytes(size_t n, size_t elem);
template <class T>
int chol_factor(hipStream_t s, size_t n, T *a, size_t lda, int64_t *info,
                void *work);
int chol_factor_lowp(hipStream_t s, int prec, size_t n, float *a, size_t lda,
                     int64_t *info, void *work);
// X = L^-T L^-1 B for nrhs vectors; x may alias b; `tmp` holds n * nrhs
// elements.
template <class T>
int chol_solve(hipStream_t s, size_t n, const T *l, size_t lda, size_t nrhs,
               const T *b, size_t ldb, T *x, size_t ldx, T *tmp);

// ------------------------------------------------------------------ triangular
// Solves on nrhs vectors: the right-hand sides come in `in` (vector r at
// in + r*ldi, overwritten as the running residual) and the solution goes to
// `out` (vector r at out + r*ldo); `in` and `out` must not overlap.
//   lower:  L y = b   (unit: diagonal taken as 1)
//   upper:  U y = b
//   lower_t: L^T y = b
template <class T>
int trsv_lower(hipStream_t s, size_t n, const T *l, size_t lda, bool unit, size_t nrhs,
               T *in, size_t ldi, T *out, size_t ldo);
template <class T>
int trsv_upper(hipStream_t s, size_t n, const T *u, size_t lda, size_t nrhs, T *in,
               size_t ldi, T *out, size_t ldo);
template <class T>
int trsv_lower_t(hipStream_t s, size_t n, const T *l, size_t lda, size_t nrhs, T *in,
                 size_t ldi, T *out, size_t ldo);
// *flag = 1 when some |m_ii| < eps (left untouched otherwise; the caller
// clears it first).
template <class T>
int diag_tiny(hipStream_t s, size_t n, const T *m, size_t lda, T eps, uint32_t *flag);

}  // namespace rlg
