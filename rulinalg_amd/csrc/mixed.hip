// Mixed-precision refined solve: factor a scaled copy of A in FP32 (or with
// FP16/BF16 trailing updates), then refine in FP64:
//
//   Dr, Dc  power-of-two row/column scalings (equilibration in the manner of
//           LAPACK dgeequ; powers of two keep the scaling exact in FP64)
//   As      = fl32(Dr A Dc), factored P As = L U in low precision
//   x0 = 0, r0 = b
//   loop:   d = U^-1 L^-1 P fl32(Dr r)       (low-precision solves)
//           x = x + Dc d                       (FP64)
//           r = b - A x                        (FP64, A read once per step)
//           eta = ||r||inf / (||A||inf ||x||inf + ||b||inf)
//   until eta <= tol (converged) or eta fails to halve (stall) or max_iter.
// On a stall the plain FP64 factorization solves the system (fallback).
//
// GMRES-IR (Carson and Higham, SIAM J. Sci. Comput. 40 (2018)) replaces the
// correction d = M^-1 r by restarted GMRES in FP64 (30 steps per cycle) on
// M^-1 A d = M^-1 r, left-preconditioned by the same low-precision factors
// widened exactly to FP64: M^-1 v = Dc U^-1 L^-1 P Dr v with FP64
// substitutions.  Factors whose classical refinement stalls then still
// converge when kappa(A) times their effective unit roundoff is not far
// above one (FP16 trailing updates at kappa 1e4 take about five outer
// steps); far beyond that (BF16 at kappa 1e4) a 30-step cycle reduces
// nothing and the stall rule falls back to FP64 as before.
#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace rlg {

// ---------------------------------------------------------------- kernels

// Max of non-negative doubles through their bit patterns (monotone for
// non-negative IEEE values).
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(addr), __double_as_longlong(v));
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double pow2_recip(double m) {
    // nearest power of two to 1/m (1 for zero rows)
    if (!(m > 0.0) || !isfinite(m)) return 1.0;
    int e;
    frexp(m, &e);  // m = f * 2^e, f in [0.5, 1)
    return ldexp(1.0, -(e - 1));
}

// Row scaling: rs[i] = 2^-round(log2 max_j |a_ij|); also anorm = max_i sum_j |a_ij|.
__global__ __launch_bounds__(256) void k_row_scale(int n, const double *__restrict__ a, long lda,
                                                   double *__restrict__ rs, double *__restrict__ anorm) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *r = a + (long)row * lda;
    double mx = 0.0, sm = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double v = fabs(r[j]);
        mx = fmax(mx, v);
        sm += v;
    }
    mx = wave_max(mx);
    sm = wave_sum(sm);
    if (lane == 0) {
        rs[row] = pow2_recip(mx);
        atomic_max_nonneg(anorm, sm);
    }
}

// Column maxima of |rs_i a_ij| into cmax (bit-pattern atomics), 64 rows per block.
__global__ __launch_bounds__(256) void k_col_max(int n, const double *__restrict__ a, long lda,
                                                 const double *__restrict__ rs, double *__restrict__ cmax) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i0 = blockIdx.y * 64;
    if (j >= n) return;
    double mx = 0.0;
    const int i1 = min(n, i0 + 64);
    for (int i = i0; i < i1; ++i) mx = fmax(mx, fabs(rs[i] * a[(long)i * lda + j]));
    atomic_max_nonneg(&cmax[j], mx);
}

__global__ void k_col_scale(int n, double *__restrict__ cs) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < n) cs[j] = pow2_recip(cs[j]);
}

// As32[i][j] = fl32(rs_i * a_ij * cs_j)
__global__ __launch_bounds__(256) void k_scale_convert(int n, const double *__restrict__ a, long lda,
                                                       const double *__restrict__ rs,
                                                       const double *__restrict__ cs,
                                                       float *__restrict__ o, long ldo) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j < n) o[(long)i * ldo + j] = (float)(rs[i] * a[(long)i * lda + j] * cs[j]);
}

// R[r][i] = B[r][i] - sum_j A[i][j] X[r][j] for NR right-hand sides starting
// at r0; one wave per row; rmax[r] = max_i |R[r][i]|.
template <int NR>
__global__ __launch_bounds__(256) void k_residual(int n, const double *__restrict__ a, long lda,
                                                  const double *__restrict__ x, long ldx,
                                                  const double *__restrict__ b, long ldb,
                                                  double *__restrict__ res, long ldr, int r0, int nr,
                                                  double *__restrict__ rmax) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *ar = a + (long)row * lda;
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double v = ar[j];
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nr) acc[q] += v * x[(long)(r0 + q) * ldx + j];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        if (q >= nr) break;
        const double s = wave_sum(acc[q]);
        if (lane == 0) {
            const double rv = b[(long)(r0 + q) * ldb + row] - s;
            res[(long)(r0 + q) * ldr + row] = rv;
            atomic_max_nonneg(&rmax[r0 + q], fabs(rv));
        }
    }
}

// max_i |v[r][i]| per vector r, into out[r] (bit-pattern atomics).
__global__ __launch_bounds__(256) void k_vec_maxabs(int n, const double *__restrict__ v, long ldv,
                                                    double *__restrict__ out) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double m = i < n ? fabs(v[(long)r * ldv + i]) : 0.0;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&out[r], m);
}

// t[r][i] = fl32(rs_i * R[r][i])
__global__ void k_scale_rhs(int n, const double *__restrict__ res, long ldr,
                            const double *__restrict__ rs, float *__restrict__ t) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) t[(long)r * n + i] = (float)(rs[i] * res[(long)r * ldr + i]);
}

// X[r][j] += cs_j * d[r][j]; xmax[r] = max_j |X[r][j]|.
__global__ void k_update_x(int n, double *__restrict__ x, long ldx, const double *__restrict__ cs,
                           const float *__restrict__ d, double *__restrict__ xmax) {
    const int r = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double m = 0.0;
    if (j < n) {
        const double v = x[(long)r * ldx + j] + cs[j] * (double)d[(long)r * n + j];
        x[(long)r * ldx + j] = v;
        m = fabs(v);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&xmax[r], m);
}

__global__ void k_zero_f64(double *p, long count) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// out[r][i] = d_i * v[r][i] (FP64).
__global__ void k_scale_vec64(int n, const double *__restrict__ d, const double *__restrict__ v,
                              double *__restrict__ out) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[(long)r * n + i] = d[i] * v[(long)r * n + i];
}

// dst = (double) src, n x n (the low-precision factors widened exactly).
__global__ void k_widen(long count, const float *__restrict__ src, double *__restrict__ dst) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = (double)src[i];
}

// Y[r][i] = sum_j A[i][j] X[r][j] for NR vectors starting at r0; one wave per row.
template <int NR>
__global__ __launch_bounds__(256) void k_matvec_multi(int n, const double *__restrict__ a, long lda,
                                                      const double *__restrict__ x, double *__restrict__ y,
                                                      int r0, int nr) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *ar = a + (long)row * lda;
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double v = ar[j];
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nr) acc[q] += v * x[(long)(r0 + q) * n + j];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        if (q >= nr) break;
        const double s = wave_sum(acc[q]);
        if (lane == 0) y[(long)(r0 + q) * n + row] = s;
    }
}

// Sum over the workgroup (256 threads); the result is valid in thread 0.
__device__ __forceinline__ double block_sum256(double v) {
    __shared__ double part[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return part[0] + part[1] + part[2] + part[3];
}

// H[i][r] = <V_i,r , Z_r> for i < cnt; basis vector i of RHS r at
// V + (i * nrhs + r) * n.  One workgroup per (i, r).
__global__ __launch_bounds__(256) void k_basis_dots(int n, int nrhs, const double *__restrict__ V,
                                                    const double *__restrict__ Z, double *__restrict__ H) {
    const int i = blockIdx.x, r = blockIdx.y;
    const double *v = V + ((long)i * nrhs + r) * n, *z = Z + (long)r * n;
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) acc += v[k] * z[k];
    const double s = block_sum256(acc);
    if (threadIdx.x == 0) H[(long)i * nrhs + r] = s;
}

// Z_r -= sum_{i < cnt} H[i][r] V_i,r.
__global__ void k_basis_axpy(int n, int nrhs, int cnt, const double *__restrict__ V,
                             const double *__restrict__ H, double *__restrict__ Z) {
    const int r = blockIdx.y;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    double z = Z[(long)r * n + k];
    for (int i = 0; i < cnt; ++i) z -= H[(long)i * nrhs + r] * V[((long)i * nrhs + r) * n + k];
    Z[(long)r * n + k] = z;
}

// nrm[r] = ||Z_r||_2; one workgroup per r.
__global__ __launch_bounds__(256) void k_norm2(int n, const double *__restrict__ Z, double *__restrict__ nrm) {
    const int r = blockIdx.x;
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const double v = Z[(long)r * n + k];
        acc += v * v;
    }
    const double s = block_sum256(acc);
    if (threadIdx.x == 0) nrm[r] = sqrt(s);
}

// out_r = Z_r / nrm[r] (zero when nrm[r] is zero).
__global__ void k_normalize(int n, const double *__restrict__ Z, const double *__restrict__ nrm,
                            double *__restrict__ out) {
    const int r = blockIdx.y;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const double d = nrm[r];
    out[(long)r * n + k] = d > 0.0 ? Z[(long)r * n + k] / d : 0.0;
}

// X_r += sum_{i < cnt} Y[i][r] V_i,r; xmax[r] = max_k |X_r[k]|.
__global__ void k_basis_combine(int n, int nrhs, int cnt, const double *__restrict__ V,
                                const double *__restrict__ Y, double *__restrict__ x, long ldx,
                                double *__restrict__ xmax) {
    const int r = blockIdx.y;
    const int k = blockIdx.x * 256 + threadIdx.x;
    double m = 0.0;
    if (k < n) {
        double v = x[(long)r * ldx + k];
        for (int i = 0; i < cnt; ++i) v += Y[(long)i * nrhs + r] * V[((long)i * nrhs + r) * n + k];
        x[(long)r * ldx + k] = v;
        m = fabs(v);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&xmax[r], m);
}

// ---------------------------------------------------------------- helpers

static int zero(hipStream_t s, double *p, size_t count) {
    if (!count) return RLG_OK;
    hipLaunchKernelGGL(k_zero_f64, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, (long)count);
    RLG_LAUNCHED();
    return RLG_OK;
}

// res = b - A x for all RHS; rmax[r] = ||res_r||inf (rmax zeroed here).
static int residual(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                    const double *x, size_t ldx, const double *b, size_t ldb, double *res, size_t ldr,
                    double *rmax) {
    RLG_TRY(zero(s, rmax, nrhs));
    // algorithmic bytes: A once, plus x and b read and res written per RHS
    RLG_PROF_W("residual_f64", s, 8.0 * ((double)n * (double)n + 3.0 * (double)n * (double)nrhs));
    for (size_t r0 = 0; r0 < nrhs; r0 += 8) {
        const int nr = (int)std::min<size_t>(8, nrhs - r0);
        hipLaunchKernelGGL(k_residual<8>, dim3(ceil_div(n, 4)), dim3(256), 0, s, (int)n, a, (long)lda,
                           x, (long)ldx, b, (long)ldb, res, (long)ldr, (int)r0, nr, rmax);
        RLG_LAUNCHED();
    }
    return RLG_OK;
}

static int vec_maxabs(hipStream_t s, size_t n, size_t nrhs, const double *v, size_t ldv, double *out) {
    RLG_TRY(zero(s, out, nrhs));
    hipLaunchKernelGGL(k_vec_maxabs, dim3(ceil_div(n, 256), (unsigned)nrhs), dim3(256), 0, s, (int)n, v,
                       (long)ldv, out);
    RLG_LAUNCHED();
    return RLG_OK;
}

static double eta_of(const std::vector<double> &h, size_t nrhs, double anorm) {
    // h = [rmax(nrhs) | xmax(nrhs) | bmax(nrhs)]
    double e = 0.0;
    for (size_t r = 0; r < nrhs; ++r) {
        const double den = anorm * h[nrhs + r] + h[2 * nrhs + r];
        const double er = den > 0 ? h[r] / den : (h[r] > 0 ? INFINITY : 0.0);
        e = std::max(e, er);
    }
    return e;
}

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, s);
    }
    double stop() {
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Enqueues the norms of the normwise backward error on stream s:
// norms = [ ||r_k||inf (nrhs) | ||x_k||inf (nrhs) | ||b_k||inf (nrhs) | ||A||inf ],
// with res (nrhs * n) and rs (n) as workspace.
int backward_error_enqueue(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                           const double *b, size_t ldb, const double *x, size_t ldx, double *nm,
                           double *res, double *rs) {
    RLG_TRY(zero(s, nm + 3 * nrhs, 1));
    hipLaunchKernelGGL(k_row_scale, dim3(ceil_div(n, 4)), dim3(256), 0, s, (int)n, a, (long)lda, rs,
                       nm + 3 * nrhs);
    RLG_LAUNCHED();
    RLG_TRY(residual(s, n, a, lda, nrhs, x, ldx, b, ldb, res, n, nm));
    RLG_TRY(vec_maxabs(s, n, nrhs, x, ldx, nm + nrhs));
    RLG_TRY(vec_maxabs(s, n, nrhs, b, ldb, nm + 2 * nrhs));
    return RLG_OK;
}

double backward_error_from(const double *norms_host, size_t nrhs) {
    std::vector<double> h(norms_host, norms_host + 3 * nrhs + 1);
    return eta_of(h, nrhs, h[3 * nrhs]);
}

int backward_error_dev(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                       const double *b, size_t ldb, const double *x, size_t ldx, double *eta) {
    PoolBuf res, norms, rs;
    RLG_TRY(res.alloc(nrhs * n * 8, s));
    RLG_TRY(norms.alloc((3 * nrhs + 1) * 8, s));
    RLG_TRY(rs.alloc(n * 8, s));
    RLG_TRY(backward_error_enqueue(s, n, a, lda, nrhs, b, ldb, x, ldx, norms.as<double>(),
                                   res.as<double>(), rs.as<double>()));
    std::vector<double> h(3 * nrhs + 1);
    RLG_HIP(hipMemcpyAsync(h.data(), norms.p, h.size() * 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    *eta = backward_error_from(h.data(), nrhs);
    return RLG_OK;
}

// ---------------------------------------------------------------- GMRES-IR

constexpr int GM_MAX = 30;             // GMRES steps per outer refinement step
constexpr double GM_INNER_TOL = 1e-6;  // relative residual of the preconditioned system

// Scratch and operands of one GMRES-IR run (FP64).
struct GmresCtx {
    hipStream_t s;
    size_t n, nrhs;
    const double *a;
    size_t lda;
    const double *rs, *cs;   // row and column scalings of the factored copy
    const double *lu;        // the low-precision factors widened to FP64 (ld n)
    const int64_t *raw;      // their row permutation
    double *t, *tmp;         // nrhs x n
    double *V;               // (GM_MAX + 1) x nrhs x n: the Krylov bases
    double *W, *Z;           // nrhs x n
    double *H;               // 2 x (GM_MAX + 1) x nrhs: one column of dot products per pass
    double *nrm;             // nrhs
    double *Y;               // GM_MAX x nrhs: the correction's coordinates
};

// out = M^-1 v = Dc U^-1 L^-1 P Dr v for the nrhs vectors of v (ld n), the
// substitutions in FP64 on the widened factors: GMRES-IR needs the
// preconditioned product in the working precision (Carson and Higham), and
// substitutions in the factors' own precision would bound the inner
// solve's accuracy by it.
static int precond(const GmresCtx &c, const double *v, double *out) {
    hipLaunchKernelGGL(k_scale_vec64, dim3(ceil_div(c.n, 256), (unsigned)c.nrhs), dim3(256), 0, c.s, (int)c.n,
                       c.rs, v, c.t);
    RLG_LAUNCHED();
    RLG_TRY(lu_solve<double>(c.s, c.n, c.lu, c.n, c.raw, c.nrhs, c.t, c.n, c.t, c.n, c.tmp));
    hipLaunchKernelGGL(k_scale_vec64, dim3(ceil_div(c.n, 256), (unsigned)c.nrhs), dim3(256), 0, c.s, (int)c.n,
                       c.cs, c.t, out);
    RLG_LAUNCHED();
    return RLG_OK;
}

static int matvec_multi(const GmresCtx &c, const double *x, double *y) {
    RLG_PROF_W("gmres_matvec", c.s, 8.0 * (double)c.n * (double)c.n);
    for (size_t r0 = 0; r0 < c.nrhs; r0 += 16) {
        const int nr = (int)std::min<size_t>(16, c.nrhs - r0);
        hipLaunchKernelGGL(k_matvec_multi<16>, dim3(ceil_div(c.n, 4)), dim3(256), 0, c.s, (int)c.n, c.a,
                           (long)c.lda, x, y, (int)r0, nr);
        RLG_LAUNCHED();
    }
    return RLG_OK;
}

// One GMRES cycle on M^-1 A d = M^-1 r for every right-hand side (r in res,
// ld n), adding d to x (ld ldx); xmax receives max |x|.  Returns the number
// of steps taken in *steps.
static int gmres_cycle(const GmresCtx &c, const double *res, double *x, size_t ldx, double *xmax, int *steps) {
    const size_t n = c.n, nr = c.nrhs;
    RLG_PROF("gmres", c.s);
    // V_0 = z / ||z||, z = M^-1 r
    RLG_TRY(precond(c, res, c.Z));
    hipLaunchKernelGGL(k_norm2, dim3((unsigned)nr), dim3(256), 0, c.s, (int)n, c.Z, c.nrm);
    RLG_LAUNCHED();
    hipLaunchKernelGGL(k_normalize, dim3(ceil_div(n, 256), (unsigned)nr), dim3(256), 0, c.s, (int)n, c.Z, c.nrm,
                       c.V);
    RLG_LAUNCHED();
    std::vector<double> beta(nr);
    RLG_HIP(hipMemcpyAsync(beta.data(), c.nrm, nr * 8, hipMemcpyDeviceToHost, c.s));
    RLG_HIP(hipStreamSynchronize(c.s));
    // Hessenberg columns (after the rotations), the rotations and the
    // rotated right-hand side g, per right-hand side
    std::vector<double> R((size_t)GM_MAX * GM_MAX * nr, 0.0), g((size_t)(GM_MAX + 1) * nr, 0.0);
    std::vector<double> gc((size_t)GM_MAX * nr), gs((size_t)GM_MAX * nr);
    std::vector<double> hcol(2 * (size_t)(GM_MAX + 1) * nr), hn(nr);
    for (size_t r = 0; r < nr; ++r) g[r] = beta[r];
    int j = 0;
    for (; j < GM_MAX; ++j) {
        double *vj = c.V + (size_t)j * nr * n, *vn = c.V + (size_t)(j + 1) * nr * n;
        RLG_TRY(matvec_multi(c, vj, c.W));
        RLG_TRY(precond(c, c.W, c.Z));
        // classical Gram-Schmidt, twice
        for (int pass = 0; pass < 2; ++pass) {
            double *Hp = c.H + (size_t)pass * (GM_MAX + 1) * nr;
            hipLaunchKernelGGL(k_basis_dots, dim3((unsigned)(j + 1), (unsigned)nr), dim3(256), 0, c.s, (int)n,
                               (int)nr, c.V, c.Z, Hp);
            RLG_LAUNCHED();
            hipLaunchKernelGGL(k_basis_axpy, dim3(ceil_div(n, 256), (unsigned)nr), dim3(256), 0, c.s, (int)n,
                               (int)nr, j + 1, c.V, Hp, c.Z);
            RLG_LAUNCHED();
        }
        hipLaunchKernelGGL(k_norm2, dim3((unsigned)nr), dim3(256), 0, c.s, (int)n, c.Z, c.nrm);
        RLG_LAUNCHED();
        hipLaunchKernelGGL(k_normalize, dim3(ceil_div(n, 256), (unsigned)nr), dim3(256), 0, c.s, (int)n, c.Z,
                           c.nrm, vn);
        RLG_LAUNCHED();
        RLG_HIP(hipMemcpyAsync(hcol.data(), c.H, hcol.size() * 8, hipMemcpyDeviceToHost, c.s));
        RLG_HIP(hipMemcpyAsync(hn.data(), c.nrm, nr * 8, hipMemcpyDeviceToHost, c.s));
        RLG_HIP(hipStreamSynchronize(c.s));
        double worst = 0.0;
        for (size_t r = 0; r < nr; ++r) {
            double *col = &R[((size_t)j * GM_MAX) * nr];  // column j: entries (i, r) at col[i * nr + r]
            for (int i = 0; i <= j; ++i)
                col[(size_t)i * nr + r] = hcol[(size_t)i * nr + r] + hcol[(size_t)(GM_MAX + 1 + i) * nr + r];
            // the earlier rotations, then a new one that zeroes h(j+1, j)
            for (int i = 0; i < j; ++i) {
                const double c0 = gc[(size_t)i * nr + r], s0 = gs[(size_t)i * nr + r];
                const double hi = col[(size_t)i * nr + r], hi1 = col[(size_t)(i + 1) * nr + r];
                col[(size_t)i * nr + r] = c0 * hi + s0 * hi1;
                col[(size_t)(i + 1) * nr + r] = -s0 * hi + c0 * hi1;
            }
            const double hjj = col[(size_t)j * nr + r], hj1 = hn[r];
            const double den = std::hypot(hjj, hj1);
            const double c1 = den > 0.0 ? hjj / den : 1.0, s1 = den > 0.0 ? hj1 / den : 0.0;
            gc[(size_t)j * nr + r] = c1;
            gs[(size_t)j * nr + r] = s1;
            col[(size_t)j * nr + r] = den;
            const double gj = g[(size_t)j * nr + r];
            g[(size_t)j * nr + r] = c1 * gj;
            g[(size_t)(j + 1) * nr + r] = -s1 * gj;
            const double rel = beta[r] > 0.0 ? std::fabs(g[(size_t)(j + 1) * nr + r]) / beta[r] : 0.0;
            worst = std::max(worst, rel);
        }
        if (worst <= GM_INNER_TOL) { ++j; break; }
    }
    const int m = std::min(j, GM_MAX);
    // back substitution R y = g per right-hand side
    std::vector<double> y((size_t)m * nr, 0.0);
    for (size_t r = 0; r < nr; ++r)
        for (int i = m - 1; i >= 0; --i) {
            double v = g[(size_t)i * nr + r];
            for (int k = i + 1; k < m; ++k) v -= R[((size_t)k * GM_MAX + i) * nr + r] * y[(size_t)k * nr + r];
            const double d = R[((size_t)i * GM_MAX + i) * nr + r];
            y[(size_t)i * nr + r] = d != 0.0 ? v / d : 0.0;
        }
    RLG_HIP(hipMemcpyAsync(c.Y, y.data(), y.size() * 8, hipMemcpyHostToDevice, c.s));
    RLG_TRY(zero(c.s, xmax, nr));
    hipLaunchKernelGGL(k_basis_combine, dim3(ceil_div(n, 256), (unsigned)nr), dim3(256), 0, c.s, (int)n, (int)nr, m,
                       c.V, c.Y, x, (long)ldx, xmax);
    RLG_LAUNCHED();
    RLG_HIP(hipStreamSynchronize(c.s));  // y goes out of scope
    *steps = m;
    return RLG_OK;
}

// ---------------------------------------------------------------- solver

int solve_mixed_dev(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                    const double *b, size_t ldb, double *x, size_t ldx, const rlg_mixed_opts &o,
                    rlg_mixed_report &rep) {
    rep = rlg_mixed_report{};
    if (n == 0 || nrhs == 0) { rep.converged = 1; return RLG_OK; }
    if (o.refine != RLG_REFINE_IR && o.refine != RLG_REFINE_GMRES) {
        set_error("refinement scheme %d is not available", o.refine);
        return RLG_E_INVALID_ARG;
    }
    const bool plain = o.factor_prec == RLG_PREC_F64;
    PoolBuf res, norms, rs, cs, lu32, ipiv, info, work, t32, tmp32;
    RLG_TRY(res.alloc(nrhs * n * 8, s));
    RLG_TRY(norms.alloc((3 * nrhs + 1) * 8, s));
    double *nm = norms.as<double>();
    std::vector<double> h(3 * nrhs + 1);
    double eta = INFINITY;
    int st = RLG_OK;

    if (!plain) {
        RLG_TRY(rs.alloc(n * 8, s));
        RLG_TRY(cs.alloc(n * 8, s));
        RLG_TRY(lu32.alloc(n * n * 4, s));
        RLG_TRY(ipiv.alloc(n * 8, s));
        RLG_TRY(info.alloc(8, s));
        RLG_TRY(work.alloc(lu_lowp_work_bytes(n), s));
        RLG_TRY(t32.alloc(nrhs * n * 4, s));
        RLG_TRY(tmp32.alloc(nrhs * n * 4, s));
        int64_t info_h = 0;
        int timeout = 0;
        {
            EventTimer tf(s);
            RLG_TRY(zero(s, nm + 3 * nrhs, 1));
            RLG_TRY(zero(s, cs.as<double>(), n));
            {
                RLG_PROF("equilibrate", s);
                hipLaunchKernelGGL(k_row_scale, dim3(ceil_div(n, 4)), dim3(256), 0, s, (int)n, a,
                                   (long)lda, rs.as<double>(), nm + 3 * nrhs);
                RLG_LAUNCHED();
                hipLaunchKernelGGL(k_col_max, dim3(ceil_div(n, 256), ceil_div(n, 64)), dim3(256), 0, s,
                                   (int)n, a, (long)lda, rs.as<double>(), cs.as<double>());
                RLG_LAUNCHED();
                hipLaunchKernelGGL(k_col_scale, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n,
                                   cs.as<double>());
                RLG_LAUNCHED();
                hipLaunchKernelGGL(k_scale_convert, dim3(ceil_div(n, 256), (unsigned)n), dim3(256), 0,
                                   s, (int)n, a, (long)lda, rs.as<double>(), cs.as<double>(),
                                   lu32.as<float>(), (long)n);
                RLG_LAUNCHED();
            }
            // FP32 panels (partial or tournament pivoting); trailing updates in
            // FP32, FP16 or BF16 (FP32 accumulation)
            RLG_TRY(lu_factor_lowp(s, o.factor_prec, o.pivot, n, lu32.as<float>(), n,
                                   ipiv.as<int64_t>(), info.as<int64_t>(), work.p));
            RLG_HIP(hipMemcpyAsync(&info_h, info.p, 8, hipMemcpyDeviceToHost, s));
            RLG_HIP(hipMemcpyAsync(&timeout, lu_timeout_flag(work.p, n, 4), 4, hipMemcpyDeviceToHost, s));
            rep.factor_ms = tf.stop();
        }
        if (timeout) { set_error("LU panel barrier timed out"); return RLG_E_DEVICE; }
        if (info_h == (int64_t)n) {
            std::vector<int64_t> ip(n), raw(n);
            RLG_HIP(hipMemcpyAsync(ip.data(), ipiv.p, n * 8, hipMemcpyDeviceToHost, s));
            RLG_HIP(hipStreamSynchronize(s));
            perm_from_ipiv(n, ip.data(), raw.data());
            RLG_HIP(hipMemcpyAsync(ipiv.p, raw.data(), n * 8, hipMemcpyHostToDevice, s));  // ipiv now holds raw
            RLG_HIP(hipStreamSynchronize(s));
            EventTimer tr(s);
            double prev = INFINITY;
            // x = 0, r = b
            for (size_t r = 0; r < nrhs; ++r)
                RLG_TRY(zero(s, x + r * ldx, n));
            RLG_TRY(vec_maxabs(s, n, nrhs, b, ldb, nm + 2 * nrhs));
            if (o.refine == RLG_REFINE_GMRES) {
                PoolBuf V, W, Z, H, nrm, Y, lu64, t64, tmp64;
                RLG_TRY(lu64.alloc(n * n * 8, s));
                RLG_TRY(t64.alloc(nrhs * n * 8, s));
                RLG_TRY(tmp64.alloc(nrhs * n * 8, s));
                hipLaunchKernelGGL(k_widen, dim3(ceil_div(n * n, 256)), dim3(256), 0, s, (long)(n * n),
                                   lu32.as<float>(), lu64.as<double>());
                RLG_LAUNCHED();
                RLG_TRY(V.alloc((size_t)(GM_MAX + 1) * nrhs * n * 8, s));
                RLG_TRY(W.alloc(nrhs * n * 8, s));
                RLG_TRY(Z.alloc(nrhs * n * 8, s));
                RLG_TRY(H.alloc(2 * (size_t)(GM_MAX + 1) * nrhs * 8, s));
                RLG_TRY(nrm.alloc(nrhs * 8, s));
                RLG_TRY(Y.alloc((size_t)GM_MAX * nrhs * 8, s));
                GmresCtx c{s, n, nrhs, a, lda, rs.as<double>(), cs.as<double>(), lu64.as<double>(),
                           ipiv.as<int64_t>(), t64.as<double>(), tmp64.as<double>(), V.as<double>(),
                           W.as<double>(), Z.as<double>(), H.as<double>(), nrm.as<double>(), Y.as<double>()};
                // residual of x = 0 is b: copy it so that every cycle reads res
                for (size_t r = 0; r < nrhs; ++r)
                    RLG_TRY(copy_f64(res.as<double>() + r * n, b + r * ldb, n, s));
                for (int it = 0; it < o.max_iter; ++it) {
                    int steps = 0;
                    RLG_TRY(gmres_cycle(c, res.as<double>(), x, ldx, nm + nrhs, &steps));
                    rep.inner_iterations += steps;
                    RLG_TRY(residual(s, n, a, lda, nrhs, x, ldx, b, ldb, res.as<double>(), n, nm));
                    RLG_HIP(hipMemcpyAsync(h.data(), nm, h.size() * 8, hipMemcpyDeviceToHost, s));
                    RLG_HIP(hipStreamSynchronize(s));
                    eta = eta_of(h, nrhs, h[3 * nrhs]);
                    rep.iterations = it + 1;
                    if (eta <= o.tol) break;
                    if (it >= 1 && !(eta <= 0.5 * prev)) break;  // stalled
                    prev = eta;
                }
            }
            const double *rcur = b;
            size_t ldrc = ldb;
            for (int it = 0; o.refine == RLG_REFINE_IR && it < o.max_iter; ++it) {
                {
                    RLG_PROF("refine_update", s);
                    hipLaunchKernelGGL(k_scale_rhs, dim3(ceil_div(n, 256), (unsigned)nrhs), dim3(256), 0, s,
                                       (int)n, rcur, (long)ldrc, rs.as<double>(), t32.as<float>());
                    RLG_LAUNCHED();
                }
                RLG_TRY(lu_solve<float>(s, n, lu32.as<float>(), n, ipiv.as<int64_t>(), nrhs,
                                        t32.as<float>(), n, t32.as<float>(), n, tmp32.as<float>()));
                RLG_TRY(zero(s, nm + nrhs, nrhs));
                {
                    RLG_PROF("refine_update", s);
                    hipLaunchKernelGGL(k_update_x, dim3(ceil_div(n, 256), (unsigned)nrhs), dim3(256), 0, s,
                                       (int)n, x, (long)ldx, cs.as<double>(), t32.as<float>(), nm + nrhs);
                    RLG_LAUNCHED();
                }
                RLG_TRY(residual(s, n, a, lda, nrhs, x, ldx, b, ldb, res.as<double>(), n, nm));
                rcur = res.as<double>();
                ldrc = n;
                RLG_HIP(hipMemcpyAsync(h.data(), nm, h.size() * 8, hipMemcpyDeviceToHost, s));
                RLG_HIP(hipStreamSynchronize(s));
                eta = eta_of(h, nrhs, h[3 * nrhs]);
                rep.iterations = it + 1;
                if (eta <= o.tol) break;
                if (it >= 1 && !(eta <= 0.5 * prev)) break;  // stalled
                prev = eta;
            }
            rep.refine_ms = tr.stop();
        }
        rep.backward_error = eta;
        rep.converged = eta <= o.tol;
        if (rep.converged) return RLG_OK;
        if (!o.fallback) {
            set_error("refinement stalled at backward error %.3e (tol %.1e)", eta, o.tol);
            return RLG_E_NOT_CONVERGED;
        }
    }

    // plain FP64 path (the drop-in factorization) -- also the fallback
    {
        EventTimer tb(s);
        PoolBuf af, ip, inf, wk, tmp;
        RLG_TRY(af.alloc(n * n * 8, s));
        RLG_TRY(ip.alloc(n * 8, s));
        RLG_TRY(inf.alloc(8, s));
        RLG_TRY(wk.alloc(lu_work_bytes(n, 8), s));
        RLG_TRY(tmp.alloc(nrhs * n * 8, s));
        RLG_HIP(hipMemcpy2DAsync(af.p, n * 8, a, lda * 8, n * 8, n, hipMemcpyDeviceToDevice, s));
        RLG_TRY(lu_factor<double>(s, n, af.as<double>(), n, ip.as<int64_t>(), inf.as<int64_t>(), wk.p));
        int64_t info_h = 0;
        int timeout = 0;
        RLG_HIP(hipMemcpyAsync(&info_h, inf.p, 8, hipMemcpyDeviceToHost, s));
        RLG_HIP(hipMemcpyAsync(&timeout, lu_timeout_flag(wk.p, n, 8), 4, hipMemcpyDeviceToHost, s));
        RLG_HIP(hipStreamSynchronize(s));
        if (timeout) { set_error("LU panel barrier timed out"); return RLG_E_DEVICE; }
        if (info_h < (int64_t)n) {
            set_error("The matrix is too ill-conditioned for LU decomposition with partial pivoting");
            return RLG_E_DIV_BY_ZERO;
        }
        std::vector<int64_t> iph(n), raw(n);
        RLG_HIP(hipMemcpyAsync(iph.data(), ip.p, n * 8, hipMemcpyDeviceToHost, s));
        RLG_HIP(hipStreamSynchronize(s));
        perm_from_ipiv(n, iph.data(), raw.data());
        RLG_HIP(hipMemcpyAsync(ip.p, raw.data(), n * 8, hipMemcpyHostToDevice, s));
        RLG_HIP(hipStreamSynchronize(s));
        RLG_TRY(lu_solve<double>(s, n, af.as<double>(), n, ip.as<int64_t>(), nrhs, b, ldb, x, ldx,
                                 tmp.as<double>()));
        st = backward_error_dev(s, n, a, lda, nrhs, b, ldb, x, ldx, &eta);
        if (plain) rep.factor_ms = tb.stop(); else rep.fallback_ms = tb.stop();
    }
    rep.fell_back = plain ? 0 : 1;
    rep.backward_error = eta;
    rep.converged = eta <= o.tol;
    return st;
}

}  // namespace rlg

using namespace rlg;

extern "C" {

void rlg_mixed_opts_default(rlg_mixed_opts *o) {
    if (!o) return;
    o->factor_prec = RLG_PREC_F16X3;
    o->refine = RLG_REFINE_IR;
    o->tol = 1e-13;
    o->max_iter = 30;
    o->fallback = 1;
    o->pivot = RLG_PIVOT_AUTO;
}

int rlg_solve_mixed_f64_dev(size_t n, const double *a, size_t lda, size_t nrhs, const double *b,
                            size_t ldb, double *x, size_t ldx, const rlg_mixed_opts *opts,
                            rlg_mixed_report *rep, void *stream) {
    if (lda < n || ldb < n || ldx < n) { set_error("bad leading dimension"); return RLG_E_INVALID_ARG; }
    rlg_mixed_opts o;
    if (opts) o = *opts; else rlg_mixed_opts_default(&o);
    if (o.factor_prec < RLG_PREC_F64 || o.factor_prec > RLG_PREC_F16X3) {
        set_error("unknown factor precision %d", o.factor_prec);
        return RLG_E_INVALID_ARG;
    }
    if (o.pivot != RLG_PIVOT_PARTIAL && o.pivot != RLG_PIVOT_TOURNAMENT && o.pivot != RLG_PIVOT_AUTO) {
        set_error("unknown pivoting %d", o.pivot);
        return RLG_E_INVALID_ARG;
    }
    if (o.pivot == RLG_PIVOT_AUTO)  // tournament where the trailing updates keep FP32 accuracy
        o.pivot = (o.factor_prec == RLG_PREC_F32 || o.factor_prec == RLG_PREC_F16X3) ? RLG_PIVOT_TOURNAMENT
                                                                                       : RLG_PIVOT_PARTIAL;
    Device *d;
    RLG_TRY(current_device(&d));
    rlg_mixed_report r;
    int st = solve_mixed_dev(pick_stream(d, stream), n, a, lda, nrhs, b, ldb, x, ldx, o, r);
    if (rep) *rep = r;
    return st;
}

int rlg_solve_mixed_f64(size_t n, const double *a, size_t lda, size_t nrhs, const double *b,
                        size_t ldb, double *x, size_t ldx, const rlg_mixed_opts *opts,
                        rlg_mixed_report *rep) {
    if (lda < n || ldb < n || ldx < n) { set_error("bad leading dimension"); return RLG_E_INVALID_ARG; }
    if (n == 0 || nrhs == 0) {
        if (rep) { *rep = rlg_mixed_report{}; rep->converged = 1; }
        return RLG_OK;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    DevBuf da, db, dx;
    RLG_TRY(da.alloc(n * n * 8));
    RLG_TRY(db.alloc(nrhs * n * 8));
    RLG_TRY(dx.alloc(nrhs * n * 8));
    RLG_HIP(hipMemcpy2DAsync(da.p, n * 8, a, lda * 8, n * 8, n, hipMemcpyHostToDevice, s));
    RLG_HIP(hipMemcpy2DAsync(db.p, n * 8, b, ldb * 8, n * 8, nrhs, hipMemcpyHostToDevice, s));
    int st = rlg_solve_mixed_f64_dev(n, da.as<double>(), n, nrhs, db.as<double>(), n, dx.as<double>(),
                                     n, opts, rep, s);
    if (st != RLG_OK && st != RLG_E_NOT_CONVERGED) return st;
    RLG_HIP(hipMemcpy2DAsync(x, ldx * 8, dx.p, n * 8, n * 8, nrhs, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    return st;
}

int rlg_backward_error_f64_dev(size_t n, const double *a, size_t lda, size_t nrhs, const double *b,
                               size_t ldb, const double *x, size_t ldx, double *eta, void *stream) {
    if (!eta) { set_error("null output"); return RLG_E_INVALID_ARG; }
    if (n == 0 || nrhs == 0) { *eta = 0.0; return RLG_OK; }
    Device *d;
    RLG_TRY(current_device(&d));
    return backward_error_dev(pick_stream(d, stream), n, a, lda, nrhs, b, ldb, x, ldx, eta);
}

}  // extern "C"
