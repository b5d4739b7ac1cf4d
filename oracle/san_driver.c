/*
 * Sanitizer driver for the CPU oracle (test infrastructure, like the oracle
 * itself).  Built with -fsanitize=address,undefined by `make -C oracle
 * sanitize`; it runs every oracle routine on exactly-sized heap buffers at
 * the edge orders (0, 1, 2, 7, 8, 9, 65) so that an out-of-bounds index or
 * undefined behaviour in a restatement is reported by the sanitizers, and
 * checks a few identities so that a silently wrong run also fails.
 * Exit status 0 = clean.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* the oracle's entry points (rlg_oracle.c) */
void orc_perm_inverse(size_t n, const size_t *perm, size_t *inv);
int orc_perm_parity(size_t n, const size_t *perm_in);
double orc_dot_f64(const double *u, const double *v, size_t len);
int orc_matmul_f64(size_t p, size_t q, size_t q2, size_t r, const double *a, size_t lda, const double *b,
                   size_t ldb, double *c);
void orc_matvec_f64(size_t rows, size_t cols, const double *a, const double *v, double *y);
int orc_lu_decompose_f64(size_t n, double *lu, size_t *perm, int64_t *ipiv, double *margin,
                         size_t *fail_step);
int orc_lu_solve_f64(size_t n, const double *lu, const size_t *perm, const double *b, double *x);
double orc_lu_det_f64(size_t n, const double *lu, const size_t *perm);
int orc_lu_inverse_f64(size_t n, const double *lu, const size_t *perm, double *inv);
void orc_lu_unpack_f64(size_t n, const double *lu, double *l, double *u);
int orc_chol_decompose_f64(size_t n, double *a, size_t *fail_col);
int orc_chol_solve_f64(size_t n, const double *l, const double *b, double *x);
double orc_chol_det_f64(size_t n, const double *l);
int orc_chol_inverse_f64(size_t n, const double *l, double *inv);
void orc_chol_unpack_f64(size_t n, double *l);
int orc_fplu_decompose_f64(size_t n, double *lu, size_t *p, size_t *q, int64_t *prow, int64_t *pcol,
                           size_t *steps);
int orc_fplu_solve_f64(size_t n, const double *lu, const size_t *p, const size_t *q, const double *b,
                       double *x);
size_t orc_fplu_rank_f64(size_t n, const double *lu);
int orc_fplu_is_invertible_f64(size_t n, const double *lu);
double orc_fplu_det_f64(size_t n, const double *lu, const size_t *p, const size_t *q);
int orc_fplu_inverse_f64(size_t n, const double *lu, const size_t *p, const size_t *q, double *inv);
int orc_lu_decompose_f32(size_t n, float *lu, size_t *perm, int64_t *ipiv, float *margin, size_t *fail_step);
int orc_chol_decompose_f32(size_t n, float *a, size_t *fail_col);
void orc_gemm_blocked_f64(size_t m, size_t k, size_t n, const double *a, size_t lda, const double *b,
                          size_t ldb, double *c);

static int failures = 0;
#define CHECK(c, ...)                                         \
    do {                                                      \
        if (!(c)) {                                           \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                     \
            fputc('\n', stderr);                              \
            ++failures;                                       \
        }                                                     \
    } while (0)

static uint64_t g_state = 88172645463325252ull;
static double rnd(void) {  /* xorshift64*, uniform in [-1, 1) */
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (double)((g_state * 2685821657736338717ull) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

/* malloc of exactly n elements (at least one byte, so n = 0 is a valid pointer
 * that any access overruns) */
static void *xalloc(size_t n, size_t elem) {
    void *p = malloc(n ? n * elem : 1);
    if (!p) abort();
    return p;
}

static double resid(size_t n, const double *a, const double *x, const double *b) {
    double r = 0, an = 0, xn = 0, bn = 0;
    for (size_t i = 0; i < n; ++i) {
        double s = 0, rs = 0;
        for (size_t j = 0; j < n; ++j) { s += a[i * n + j] * x[j]; rs += fabs(a[i * n + j]); }
        r = fmax(r, fabs(b[i] - s));
        an = fmax(an, rs);
        xn = fmax(xn, fabs(x[i]));
        bn = fmax(bn, fabs(b[i]));
    }
    return an * xn + bn > 0 ? r / (an * xn + bn) : r;
}

static void run_order(size_t n) {
    double *a = xalloc(n * n, 8), *w = xalloc(n * n, 8), *inv = xalloc(n * n, 8);
    double *l = xalloc(n * n, 8), *u = xalloc(n * n, 8), *s = xalloc(n * n, 8);
    double *b = xalloc(n, 8), *x = xalloc(n, 8), *y = xalloc(n, 8), *margin = xalloc(n, 8);
    size_t *perm = xalloc(n, sizeof(size_t)), *pinv = xalloc(n, sizeof(size_t)), *q = xalloc(n, sizeof(size_t));
    int64_t *ipiv = xalloc(n, 8), *pcol = xalloc(n, 8);
    size_t fail = 0, steps = 0;
    for (size_t i = 0; i < n * n; ++i) a[i] = rnd();
    for (size_t i = 0; i < n; ++i) b[i] = rnd();

    /* products */
    CHECK(orc_matmul_f64(n, n, n, n, a, n, a, n, w) == 0, "matmul n=%zu", n);
    orc_gemm_blocked_f64(n, n, n, a, n, a, n, inv);
    for (size_t i = 0; i < n * n; ++i) CHECK(fabs(w[i] - inv[i]) <= 1e-12 * (double)(n + 1), "gemm n=%zu", n);
    orc_matvec_f64(n, n, a, b, y);
    (void)orc_dot_f64(a, b, n);

    /* PartialPivLu */
    memcpy(w, a, n * n * 8);
    CHECK(orc_lu_decompose_f64(n, w, perm, ipiv, margin, &fail) == 0, "lu n=%zu", n);
    CHECK(orc_lu_solve_f64(n, w, perm, b, x) == 0, "lu solve n=%zu", n);
    CHECK(resid(n, a, x, b) < 1e-13, "lu residual n=%zu", n);
    (void)orc_lu_det_f64(n, w, perm);
    CHECK(orc_lu_inverse_f64(n, w, perm, inv) == 0, "lu inverse n=%zu", n);
    orc_lu_unpack_f64(n, w, l, u);
    orc_perm_inverse(n, perm, pinv);
    (void)orc_perm_parity(n, perm);

    /* Cholesky on S = A A^T / n + I */
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            double t = 0;
            for (size_t k = 0; k < n; ++k) t += a[i * n + k] * a[j * n + k];
            s[i * n + j] = t / (double)n + (i == j ? 1.0 : 0.0);
        }
    memcpy(w, s, n * n * 8);
    CHECK(orc_chol_decompose_f64(n, w, &fail) == 0, "chol n=%zu", n);
    CHECK(orc_chol_solve_f64(n, w, b, x) == 0, "chol solve n=%zu", n);
    CHECK(resid(n, s, x, b) < 1e-13, "chol residual n=%zu", n);
    (void)orc_chol_det_f64(n, w);
    CHECK(orc_chol_inverse_f64(n, w, inv) == 0, "chol inverse n=%zu", n);
    orc_chol_unpack_f64(n, w);

    /* FullPivLu, regular and with a zero trailing block */
    memcpy(w, a, n * n * 8);
    CHECK(orc_fplu_decompose_f64(n, w, perm, q, ipiv, pcol, &steps) == 0 && steps == n, "fplu n=%zu", n);
    CHECK(orc_fplu_solve_f64(n, w, perm, q, b, x) == 0, "fplu solve n=%zu", n);
    (void)orc_fplu_rank_f64(n, w);
    (void)orc_fplu_is_invertible_f64(n, w);
    (void)orc_fplu_det_f64(n, w, perm, q);
    CHECK(orc_fplu_inverse_f64(n, w, perm, q, inv) == 0, "fplu inverse n=%zu", n);
    for (size_t i = 0; i < n * n; ++i) w[i] = 0.0;
    if (n) w[0] = 1.0;
    CHECK(orc_fplu_decompose_f64(n, w, perm, q, ipiv, pcol, &steps) == 0 && steps == (n ? 1 : 0),
          "fplu rank-one n=%zu", n);
    CHECK(orc_fplu_rank_f64(n, w) == (n ? 1u : 0u), "fplu rank n=%zu", n);

    /* f32 forms */
    {
        float *wf = xalloc(n * n, 4), *mf = xalloc(n, 4);
        for (size_t i = 0; i < n * n; ++i) wf[i] = (float)a[i];
        (void)orc_lu_decompose_f32(n, wf, perm, ipiv, mf, &fail);
        for (size_t i = 0; i < n * n; ++i) wf[i] = (float)s[i];
        (void)orc_chol_decompose_f32(n, wf, &fail);
        free(wf);
        free(mf);
    }
    free(a); free(w); free(inv); free(l); free(u); free(s); free(b); free(x); free(y); free(margin);
    free(perm); free(pinv); free(q); free(ipiv); free(pcol);
}

int main(void) {
    const size_t orders[] = {0, 1, 2, 7, 8, 9, 16, 65};
    for (size_t i = 0; i < sizeof orders / sizeof orders[0]; ++i) run_order(orders[i]);
    /* failure paths: a singular LU and an indefinite Cholesky */
    {
        double z[9] = {0, 1, 2, 0, 3, 4, 0, 5, 6};
        size_t perm[3], fail = 99;
        int64_t ipiv[3];
        CHECK(orc_lu_decompose_f64(3, z, perm, ipiv, NULL, &fail) == 4 && fail == 0, "singular lu");
        double m[4] = {1, 2, 2, 1};
        CHECK(orc_chol_decompose_f64(2, m, &fail) == 2 && fail == 1, "indefinite chol");
    }
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("oracle sanitizer run: clean\n");
    return 0;
}
