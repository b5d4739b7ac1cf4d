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

// ---------------------------------------------------------------- solver

int solve_mixed_dev(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                    const double *b, size_t ldb, double *x, size_t ldx, const rlg_mixed_opts &o,
                    rlg_mixed_report &rep) {
    rep = rlg_mixed_report{};
    if (n == 0 || nrhs == 0) { rep.converged = 1; return RLG_OK; }
    if (o.refine != RLG_REFINE_IR) {
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
            const double *rcur = b;
            size_t ldrc = ldb;
            for (int it = 0; it < o.max_iter; ++it) {
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
    o->pivot = RLG_PIVOT_PARTIAL;
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
    if (o.pivot != RLG_PIVOT_PARTIAL && o.pivot != RLG_PIVOT_TOURNAMENT) {
        set_error("unknown pivoting %d", o.pivot);
        return RLG_E_INVALID_ARG;
    }
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
