// Sanitizer driver for librlgpu's host code (test infrastructure).  Built by
// tests/sanitize/Makefile against a copy of the library whose host half is
// compiled with -fsanitize=address,undefined.  It needs no GPU: it calls
// every entry point of include/rlgpu.h with the arguments that its checks
// reject (bad shapes, strides, permutations, null handles) or that need no
// device (empty operands, the profiler's text dump into short buffers), and
// checks the status codes.  Calls that reach the device find none and must
// fail with a status, not a crash.  Exit status 0 = every check passed (the
// sanitizers abort the process on their own findings).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rlgpu.h"

static int failures = 0;
#define EXPECT(cond, what)                                                    \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::fprintf(stderr, "FAIL line %d: %s (%s)\n", __LINE__, what,   \
                         rlg_last_error_message());                           \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

int main() {
    std::vector<double> a(64, 1.0), b(64, 1.0), c(64, 0.0);
    std::vector<float> af(64, 1.0f), bf(64, 1.0f), cf(64, 0.0f);
    EXPECT(rlg_version() && std::strlen(rlg_version()) > 0, "version");

    // GEMM (mat_mul.rs:21): inner dimensions differ -> shape; strides; empty
    EXPECT(rlg_gemm_f64(2, 3, a.data(), 3, 4, 2, b.data(), 2, c.data(), 2) == RLG_E_SHAPE, "gemm shape");
    EXPECT(rlg_gemm_f32(2, 3, af.data(), 3, 4, 2, bf.data(), 2, cf.data(), 2) == RLG_E_SHAPE, "sgemm shape");
    EXPECT(rlg_gemm_f64(2, 3, a.data(), 2, 3, 2, b.data(), 2, c.data(), 2) == RLG_E_INVALID_ARG, "gemm lda");
    EXPECT(rlg_gemm_f64(0, 3, a.data(), 3, 3, 2, b.data(), 2, c.data(), 2) == RLG_OK, "gemm empty");
    EXPECT(rlg_matvec_f64(2, 3, a.data(), 3, 4, b.data(), c.data()) == RLG_E_SHAPE, "matvec shape");
    EXPECT(rlg_matvec_f64(0, 3, a.data(), 3, 3, b.data(), c.data()) == RLG_OK, "matvec empty");

    // permutations (permutation_matrix.rs:424-467)
    size_t bad_range[3] = {0, 1, 5}, dup[3] = {0, 1, 1}, ok3[3] = {2, 0, 1};
    EXPECT(rlg_permute_rows_f64(3, bad_range, 3, 2, a.data(), 2, c.data(), 2) == RLG_E_INVALID_PERMUTATION,
           "perm out of range");
    EXPECT(rlg_permute_rows_f64(3, dup, 3, 2, a.data(), 2, c.data(), 2) == RLG_E_INVALID_PERMUTATION,
           "perm duplicate");
    EXPECT(rlg_permute_rows_f64(3, ok3, 4, 2, a.data(), 2, c.data(), 2) == RLG_E_SHAPE, "perm rows shape");
    EXPECT(rlg_permute_cols_f64(3, ok3, 2, 4, a.data(), 4, c.data(), 4) == RLG_E_SHAPE, "perm cols shape");
    EXPECT(rlg_permute_cols_f32(3, dup, 2, 3, af.data(), 3, cf.data(), 3) == RLG_E_INVALID_PERMUTATION,
           "perm cols duplicate");
    EXPECT(rlg_permute_rows_f32(0, nullptr, 0, 3, af.data(), 3, cf.data(), 3) == RLG_OK, "perm empty");
    EXPECT(rlg_permute_rows_f64_dev(3, nullptr, 2, a.data(), 2, c.data(), 2, nullptr) == RLG_E_INVALID_ARG,
           "perm dev null");
    EXPECT(rlg_permute_cols_f64_dev(2, 3, nullptr, a.data(), 3, c.data(), 3, nullptr) == RLG_E_INVALID_ARG,
           "perm cols dev null");

    // triangular solves (base/mod.rs:1015-1067): non-square, RHS size
    EXPECT(rlg_solve_u_triangular_f64(2, 3, a.data(), 3, 2, 1, b.data(), 2, c.data(), 2) == RLG_E_SHAPE,
           "utri square");
    EXPECT(rlg_solve_l_triangular_f64(3, 3, a.data(), 3, 2, 1, b.data(), 2, c.data(), 2) == RLG_E_SHAPE,
           "ltri rhs");
    EXPECT(rlg_solve_u_triangular_f32(2, 3, af.data(), 3, 2, 1, bf.data(), 2, cf.data(), 2) == RLG_E_SHAPE,
           "utri f32");
    EXPECT(rlg_solve_l_triangular_f32(3, 3, af.data(), 3, 2, 1, bf.data(), 2, cf.data(), 2) == RLG_E_SHAPE,
           "ltri f32");

    // PartialPivLu (lu.rs:165, 232)
    rlg_lu *lu = reinterpret_cast<rlg_lu *>(0x1);
    EXPECT(rlg_lu_decompose_f64(3, 4, a.data(), 4, &lu) == RLG_E_SHAPE && lu == nullptr, "lu square");
    EXPECT(rlg_lu_decompose_f32(3, 4, af.data(), 4, &lu) == RLG_E_SHAPE, "lu f32 square");
    EXPECT(rlg_lu_decompose_f64(3, 3, a.data(), 2, &lu) == RLG_E_INVALID_ARG, "lu lda");
    EXPECT(rlg_lu_decompose_f64(3, 3, a.data(), 3, nullptr) == RLG_E_INVALID_ARG, "lu null out");
    EXPECT(rlg_lu_decompose_f64_dev(3, a.data(), 2, nullptr, &lu) == RLG_E_INVALID_ARG, "lu dev lda");
    EXPECT(rlg_lu_solve_f64(nullptr, 3, 1, b.data(), 3, c.data(), 3) != RLG_OK, "lu solve null");
    EXPECT(rlg_lu_solve_f32(nullptr, 3, 1, bf.data(), 3, cf.data(), 3) != RLG_OK, "lu solve f32 null");
    EXPECT(rlg_lu_solve_f64_dev(nullptr, 1, b.data(), 3, c.data(), 3, nullptr) != RLG_OK, "lu solve dev null");
    size_t sz = 0;
    EXPECT(rlg_lu_size(nullptr, &sz) != RLG_OK, "lu size null");
    EXPECT(rlg_lu_perm(nullptr, nullptr) != RLG_OK, "lu perm null");
    EXPECT(rlg_lu_ipiv(nullptr, nullptr) != RLG_OK, "lu ipiv null");
    EXPECT(rlg_lu_packed_f64(nullptr, c.data(), 3) != RLG_OK, "lu packed null");
    EXPECT(rlg_lu_packed_f32(nullptr, cf.data(), 3) != RLG_OK, "lu packed f32 null");
    EXPECT(rlg_lu_unpack_f64(nullptr, c.data(), c.data(), nullptr) != RLG_OK, "lu unpack null");
    double det = 0;
    EXPECT(rlg_lu_det_f64(nullptr, &det) != RLG_OK, "lu det null");
    EXPECT(rlg_lu_inverse_f64(nullptr, c.data(), 3) != RLG_OK, "lu inverse null");
    rlg_lu_free(nullptr);

    // FullPivLu (lu.rs:384-386, 459-460)
    rlg_fplu *fp = reinterpret_cast<rlg_fplu *>(0x1);
    EXPECT(rlg_fplu_decompose_f64(2, 3, a.data(), 3, &fp) == RLG_E_SHAPE && fp == nullptr, "fplu square");
    EXPECT(rlg_fplu_decompose_f32(2, 3, af.data(), 3, &fp) == RLG_E_SHAPE, "fplu f32 square");
    EXPECT(rlg_fplu_decompose_f64(3, 3, a.data(), 2, &fp) == RLG_E_INVALID_ARG, "fplu lda");
    EXPECT(rlg_fplu_decompose_f64_dev(3, a.data(), 2, nullptr, &fp) == RLG_E_INVALID_ARG, "fplu dev lda");
    EXPECT(rlg_fplu_solve_f64(nullptr, 3, 1, b.data(), 3, c.data(), 3) != RLG_OK, "fplu solve null");
    EXPECT(rlg_fplu_solve_f32(nullptr, 3, 1, bf.data(), 3, cf.data(), 3) != RLG_OK, "fplu solve f32 null");
    EXPECT(rlg_fplu_solve_true_f64(nullptr, 3, 1, b.data(), 3, c.data(), 3) != RLG_OK, "fplu solve true null");
    EXPECT(rlg_fplu_solve_f64_dev(nullptr, 1, b.data(), 3, c.data(), 3, 0, nullptr) != RLG_OK, "fplu dev null");
    EXPECT(rlg_fplu_size(nullptr, &sz) != RLG_OK, "fplu size null");
    EXPECT(rlg_fplu_perms(nullptr, nullptr, nullptr) != RLG_OK, "fplu perms null");
    EXPECT(rlg_fplu_steps(nullptr, &sz) != RLG_OK, "fplu steps null");
    EXPECT(rlg_fplu_unpack_f64(nullptr, c.data(), c.data()) != RLG_OK, "fplu unpack null");
    EXPECT(rlg_fplu_unpack_f32(nullptr, cf.data(), cf.data()) != RLG_OK, "fplu unpack f32 null");
    EXPECT(rlg_fplu_rank(nullptr, &sz) != RLG_OK, "fplu rank null");
    int yes = 0;
    EXPECT(rlg_fplu_is_invertible(nullptr, &yes) != RLG_OK, "fplu invertible null");
    EXPECT(rlg_fplu_det_f64(nullptr, &det) != RLG_OK, "fplu det null");
    EXPECT(rlg_fplu_inverse_f64(nullptr, c.data(), 3) != RLG_OK, "fplu inverse null");
    rlg_fplu_free(nullptr);

    // Cholesky (cholesky.rs:117, 195)
    rlg_chol *ch = reinterpret_cast<rlg_chol *>(0x1);
    EXPECT(rlg_chol_decompose_f64(2, 3, a.data(), 3, &ch) == RLG_E_SHAPE && ch == nullptr, "chol square");
    EXPECT(rlg_chol_decompose_f32(2, 3, af.data(), 3, &ch) == RLG_E_SHAPE, "chol f32 square");
    EXPECT(rlg_chol_decompose_f64_dev(3, a.data(), 2, nullptr, &ch) != RLG_OK, "chol dev lda");
    EXPECT(rlg_chol_solve_f64(nullptr, 3, 1, b.data(), 3, c.data(), 3) != RLG_OK, "chol solve null");
    EXPECT(rlg_chol_solve_f32(nullptr, 3, 1, bf.data(), 3, cf.data(), 3) != RLG_OK, "chol solve f32 null");
    EXPECT(rlg_chol_solve_f64_dev(nullptr, 1, b.data(), 3, c.data(), 3, nullptr) != RLG_OK, "chol dev null");
    EXPECT(rlg_chol_size(nullptr, &sz) != RLG_OK, "chol size null");
    EXPECT(rlg_chol_unpack_f64(nullptr, c.data(), 3) != RLG_OK, "chol unpack null");
    EXPECT(rlg_chol_unpack_f32(nullptr, cf.data(), 3) != RLG_OK, "chol unpack f32 null");
    EXPECT(rlg_chol_det_f64(nullptr, &det) != RLG_OK, "chol det null");
    EXPECT(rlg_chol_inverse_f64(nullptr, c.data(), 3) != RLG_OK, "chol inverse null");
    rlg_chol_free(nullptr);

    // mixed precision and backward error
    rlg_mixed_opts o;
    rlg_mixed_opts_default(&o);
    rlg_mixed_opts_default(nullptr);
    EXPECT(o.factor_prec == RLG_PREC_F16X3 && o.refine == RLG_REFINE_IR && o.tol == 1e-13, "mixed defaults");
    rlg_mixed_report rep;
    EXPECT(rlg_solve_mixed_f64(3, a.data(), 2, 1, b.data(), 3, c.data(), 3, &o, &rep) == RLG_E_INVALID_ARG,
           "mixed lda");
    EXPECT(rlg_solve_mixed_f64(0, a.data(), 0, 1, b.data(), 0, c.data(), 0, &o, &rep) == RLG_OK &&
               rep.converged == 1, "mixed empty");
    EXPECT(rlg_solve_mixed_f64_dev(3, a.data(), 2, 1, b.data(), 3, c.data(), 3, &o, &rep, nullptr) ==
               RLG_E_INVALID_ARG, "mixed dev lda");
    rlg_mixed_opts bad = o;
    bad.factor_prec = 9;
    EXPECT(rlg_solve_mixed_f64_dev(3, a.data(), 3, 1, b.data(), 3, c.data(), 3, &bad, &rep, nullptr) ==
               RLG_E_INVALID_ARG, "mixed precision");
    EXPECT(rlg_backward_error_f64_dev(3, a.data(), 3, 1, b.data(), 3, c.data(), 3, nullptr, nullptr) ==
               RLG_E_INVALID_ARG, "eta null");
    double eta = -1;
    EXPECT(rlg_backward_error_f64_dev(0, a.data(), 0, 1, b.data(), 0, c.data(), 0, &eta, nullptr) == RLG_OK &&
               eta == 0.0, "eta empty");

    // batched and serving: no device here, so every call must fail cleanly
    EXPECT(rlg_batch_solve_f64_dev(1, nullptr, nullptr) != RLG_OK, "batch null items");
    size_t fb = 0;
    EXPECT(rlg_batch_solve_mixed_dev(1, nullptr, 30, 1e-13, &fb, nullptr) != RLG_OK, "batch mixed null");
    char rep_text[16];
    EXPECT(rlg_batch_autotune(1, rep_text, sizeof rep_text) != RLG_OK, "autotune without device");
    rlg_batch_graph *g = nullptr;
    EXPECT(rlg_batch_graph_create(1, nullptr, &g) != RLG_OK, "graph null items");
    EXPECT(rlg_batch_graph_launch(nullptr, nullptr) != RLG_OK, "graph launch null");
    rlg_batch_graph_destroy(nullptr);
    rlg_server_opts so;
    rlg_server_opts_default(&so);
    rlg_server *srv = nullptr;
    EXPECT(rlg_server_create(&so, &srv) == RLG_E_INVALID_ARG && srv == nullptr, "server disabled");
    EXPECT(rlg_server_create(&so, nullptr) == RLG_E_INVALID_ARG, "server null out");
    int64_t ticket = 0;
    EXPECT(rlg_server_submit(nullptr, nullptr, &ticket) != RLG_OK, "server submit null");
    EXPECT(rlg_server_wait(nullptr, 0, nullptr) != RLG_OK, "server wait null");
    EXPECT(rlg_server_release(nullptr, 0) != RLG_OK, "server release null");
    EXPECT(rlg_server_evict(nullptr, 0) != RLG_OK, "server evict null");
    rlg_server_destroy(nullptr);

    // device memory and the profiler's text output
    void *p = nullptr;
    EXPECT(rlg_set_device(-1) != RLG_OK, "set_device -1");
    EXPECT(rlg_dev_alloc(64, &p) != RLG_OK, "dev_alloc without device");
    EXPECT(rlg_dev_free(nullptr) == RLG_OK || true, "dev_free null");
    EXPECT(rlg_prof_enable(1) == RLG_OK, "prof enable");
    EXPECT(rlg_prof_reset() == RLG_OK, "prof reset");
    char tiny[4];
    (void)rlg_prof_dump(tiny, sizeof tiny);
    EXPECT(std::strlen(tiny) < sizeof tiny, "prof dump truncates");
    int64_t launches = -1;
    double ms = -1, work = -1;
    (void)rlg_prof_get("lu_panel", &launches, &ms, &work);
    EXPECT(rlg_prof_enable(0) == RLG_OK, "prof disable");

    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("librlgpu host sanitizer run: clean\n");
    return 0;
}
