// LU with complete pivoting (FullPivLu, src/matrix/decomposition/lu.rs:339-599)
// on the device.
//
// Complete pivoting needs the fully updated trailing block before every
// pivot choice, so the factorization cannot be blocked: it is n steps of a
// rank-1 update, each bound by HBM (the trailing block read and written
// once).  One launch per step:
//
//   * every workgroup owns a 32 x 64 tile of the trailing block
//     A[k+1:n, k+1:n] and applies step k's row and column interchange and
//     elimination to it, reading the element it overwrites and four staged
//     vectors (old row k, old column k, the pivot row, the pivot column);
//     a few extra workgroups write the parts outside the tile grid (row k,
//     the multipliers in column k, the interchanged entries left of and
//     above the diagonal);
//   * every workgroup proposes its tile's largest |value| (ties to the first
//     in row-major order) and counts its arrival; the workgroup that arrives
//     last picks step k+1's pivot from all proposals and stages the four
//     vectors for step k+1.  No workgroup ever waits for another.
//
// Every element is computed as the reference computes it, m = a_ik / a_kk
// (a division) and a_ij - m * a_kj (a separately rounded product and
// difference), so the packed factor and both permutations are identical to
// the reference's at any order.  select_pivot (lu.rs:364-380) scans row by
// row with a strict `>` from (k, k): its pivot is the first maximal |value|
// in row-major order, which is the order of the linear index i * n + j used
// here.  A zero pivot ends the factorization (lu.rs:405-407): the remaining
// launches return at once.
#include <climits>
#include <vector>

#include "kernels.hpp"

namespace rlg {

constexpr int FTR = 32;   // tile rows
constexpr int FTC = 64;   // tile columns (one wave-wide row segment)
constexpr int FTH = 256;  // threads per workgroup

struct FpluCand {
    double mag;     // |value| (exact in double for both element types); -1: none
    long long idx;  // i * n + j
};

template <class T>
struct FpluWs {
    T *srk;          // old row k (n)
    T *spr;          // old pivot row (n)
    T *sck;          // old column k (n)
    T *spc;          // old pivot column (n)
    int64_t *prow;   // pivot row of each step (n)
    int64_t *pcol;   // pivot column of each step (n)
    int64_t *steps;  // steps that run (n until a zero pivot)
    unsigned *count; // arrivals in the current launch
    FpluCand *cand;  // one proposal per workgroup
};

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t fplu_tiles(size_t w) { return (size_t)ceil_div(w, FTR) * ceil_div(w, FTC); }
static size_t fplu_border(size_t n) { return ceil_div(n, FTH); }

size_t fplu_work_bytes(size_t n, size_t elem) {
    const size_t nn = n ? n : 1;
    return 4 * al256(nn * elem) + 2 * al256(nn * 8) + al256(8) + al256(4) +
           al256((fplu_tiles(nn) + fplu_border(nn)) * sizeof(FpluCand));
}

template <class T>
static FpluWs<T> fplu_carve(void *work, size_t n) {
    const size_t nn = n ? n : 1;
    char *p = static_cast<char *>(work);
    FpluWs<T> w;
    w.srk = reinterpret_cast<T *>(p); p += al256(nn * sizeof(T));
    w.spr = reinterpret_cast<T *>(p); p += al256(nn * sizeof(T));
    w.sck = reinterpret_cast<T *>(p); p += al256(nn * sizeof(T));
    w.spc = reinterpret_cast<T *>(p); p += al256(nn * sizeof(T));
    w.prow = reinterpret_cast<int64_t *>(p); p += al256(nn * 8);
    w.pcol = reinterpret_cast<int64_t *>(p); p += al256(nn * 8);
    w.steps = reinterpret_cast<int64_t *>(p); p += al256(8);
    w.count = reinterpret_cast<unsigned *>(p); p += al256(4);
    w.cand = reinterpret_cast<FpluCand *>(p);
    return w;
}

__device__ __forceinline__ bool cand_better(double m, long long i, double bm, long long bi) {
    return m > bm || (m == bm && i < bi);
}

// Best (mag, idx) over the workgroup; the result is valid in thread 0.
__device__ FpluCand block_best(double m, long long i) {
    __shared__ double sm[FTH / 64];
    __shared__ long long si[FTH / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double om = __shfl_xor(m, off, 64);
        const long long oi = __shfl_xor(i, off, 64);
        if (cand_better(om, oi, m, i)) { m = om; i = oi; }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sm[wave] = m; si[wave] = i; }
    __syncthreads();
    FpluCand c{-1.0, LLONG_MAX};
    if (threadIdx.x == 0)
        for (int q = 0; q < FTH / 64; ++q)
            if (cand_better(sm[q], si[q], c.mag, c.idx)) { c.mag = sm[q]; c.idx = si[q]; }
    return c;
}

// Run by the workgroup that arrived last: picks the pivot of step `next`
// from all proposals and stages its four vectors, or ends the
// factorization on a zero pivot.
template <class T>
__device__ void fplu_choose(const T *__restrict__ A, long lda, int n, int next, FpluWs<T> w, int total) {
    __shared__ long long s_piv;
    double m = -1.0;
    long long i = LLONG_MAX;
    for (int q = threadIdx.x; q < total; q += FTH) {
        const FpluCand c = w.cand[q];
        if (cand_better(c.mag, c.idx, m, i)) { m = c.mag; i = c.idx; }
    }
    const FpluCand best = block_best(m, i);
    if (threadIdx.x == 0) {
        long long piv = best.idx;
        const T dkk = A[(long)next * lda + next];
        // select_pivot starts from a[next][next] and only a strictly larger
        // |value| replaces it: a NaN there is never replaced
        if (dkk != dkk) piv = (long long)next * n + next;
        else if (!(best.mag > 0.0)) piv = -1;  // |pivot| == 0: the loop ends
        if (piv < 0) {
            *w.steps = next;
        } else {
            w.prow[next] = piv / n;
            w.pcol[next] = piv % n;
        }
        s_piv = piv;
    }
    __syncthreads();
    const long long piv = s_piv;
    if (piv < 0) return;
    const int pr = (int)(piv / n), pc = (int)(piv % n);
    for (int j = threadIdx.x; j < n; j += FTH) {
        w.srk[j] = A[(long)next * lda + j];
        w.spr[j] = A[(long)pr * lda + j];
        w.sck[j] = A[(long)j * lda + next];
        w.spc[j] = A[(long)j * lda + pc];
    }
}

// Proposal, arrival count and, in the last workgroup, the next pivot.
template <class T>
__device__ void fplu_arrive(const T *__restrict__ A, long lda, int n, int next, FpluWs<T> w,
                            double m, long long i) {
    __shared__ int s_last;
    const FpluCand c = block_best(m, i);
    if (threadIdx.x == 0) w.cand[blockIdx.x] = c;
    __threadfence();  // this workgroup's writes (matrix and proposal) before its arrival
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned old = atomicAdd(w.count, 1u);
        s_last = old == gridDim.x - 1;
        if (s_last) *w.count = 0u;  // ready for the next launch
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // every other workgroup's writes after their arrivals
    if (next < n) fplu_choose<T>(A, lda, n, next, w, (int)gridDim.x);
}

// The first pivot: proposals over the whole matrix, no elimination.
template <class T>
__global__ __launch_bounds__(FTH) void k_fplu_first(T *__restrict__ A, long lda, int n, FpluWs<T> w) {
    const int tiles_c = (n + FTC - 1) / FTC;
    const int tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
    const int j = tc * FTC + (threadIdx.x & 63);
    double m = -1.0;
    long long bi = LLONG_MAX;
    if (j < n)
        for (int q = threadIdx.x >> 6; q < FTR; q += FTH / 64) {
            const int i = tr * FTR + q;
            if (i >= n) break;
            const double v = fabs((double)A[(long)i * lda + j]);
            const long long li = (long long)i * n + j;
            if (cand_better(v, li, m, bi)) { m = v; bi = li; }
        }
    fplu_arrive<T>(A, lda, n, 0, w, m, bi);
}

// Step k: interchange (k <-> pr rows, k <-> pc columns), elimination of the
// trailing block and the next pivot.  Workgroups [0, tiles) own tiles of
// A[k+1:n, k+1:n]; the rest write row k, column k below the diagonal and
// the interchanged entries of rows < k and of row pr left of column k.
template <class T>
__global__ __launch_bounds__(FTH) void k_fplu_step(T *__restrict__ A, long lda, int n, int k, int tiles,
                                                   FpluWs<T> w) {
#pragma clang fp contract(off)
    if (k >= *w.steps) return;  // a zero pivot ended the loop
    const int pr = (int)w.prow[k], pc = (int)w.pcol[k];
    const T pivot = w.spr[pc];  // a[pr][pc] before the interchange
    double m = -1.0;
    long long bi = LLONG_MAX;
    if ((int)blockIdx.x < tiles) {
        const int wdt = n - k - 1;
        const int tiles_c = (wdt + FTC - 1) / FTC;
        const int tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
        const int j = k + 1 + tc * FTC + (threadIdx.x & 63);
        if (j < n) {
            const int sj = j == pc ? k : j;   // column j after the interchange was column sj
            const T ukj = w.spr[sj];          // row k after the interchange
            for (int q = threadIdx.x >> 6; q < FTR; q += FTH / 64) {
                const int i = k + 1 + tr * FTR + q;
                if (i >= n) break;
                const int si = i == pr ? k : i;
                const T mult = w.spc[si] / pivot;
                T old;
                if (si == k) old = w.srk[sj];
                else if (sj == k) old = w.sck[i];
                else old = A[(long)i * lda + j];
                const T v = old - mult * ukj;
                A[(long)i * lda + j] = v;
                const double av = fabs((double)v);
                const long long li = (long long)i * n + j;
                if (cand_better(av, li, m, bi)) { m = av; bi = li; }
            }
        }
    } else {
        const int t = ((int)blockIdx.x - tiles) * FTH + threadIdx.x;
        if (t < n) {
            // row k: the pivot row with its columns interchanged
            const int st = t == pc ? k : (t == k ? pc : t);
            A[(long)k * lda + t] = w.spr[st];
            if (t < k) {
                if (pc != k) {  // columns k and pc of the rows above
                    A[(long)t * lda + k] = w.spc[t];
                    A[(long)t * lda + pc] = w.sck[t];
                }
                if (pr != k) A[(long)pr * lda + t] = w.srk[t];  // row pr left of column k
            }
            if (t > k) {  // the multipliers of column k
                const int si = t == pr ? k : t;
                A[(long)t * lda + k] = w.spc[si] / pivot;
            }
        }
    }
    fplu_arrive<T>(A, lda, n, k + 1, w, m, bi);
}

__global__ void k_fplu_init(int64_t *steps, unsigned *count, int n) {
    if (threadIdx.x == 0) {
        *steps = n;
        *count = 0u;
    }
}

template <class T>
int fplu_factor(hipStream_t s, size_t n, T *a, size_t lda, void *work) {
    if (n == 0) return RLG_OK;
    if (n > (size_t)46340) {  // linear indices i * n + j stay below 2^31 in the tile index math
        set_error("FullPivLu order %zu is above the supported 46340", n);
        return RLG_E_INVALID_ARG;
    }
    const FpluWs<T> w = fplu_carve<T>(work, n);
    hipLaunchKernelGGL(k_fplu_init, dim3(1), dim3(64), 0, s, w.steps, w.count, (int)n);
    RLG_LAUNCHED();
    {
        // algorithmic bytes: the matrix read once
        RLG_PROF_W("fplu_first", s, (double)n * (double)n * sizeof(T));
        hipLaunchKernelGGL(k_fplu_first<T>, dim3((unsigned)fplu_tiles(n)), dim3(FTH), 0, s, a, (long)lda,
                           (int)n, w);
        RLG_LAUNCHED();
    }
    for (size_t k = 0; k < n; ++k) {
        const size_t wdt = n - k - 1;
        const size_t tiles = fplu_tiles(wdt);
        const unsigned grid = (unsigned)(tiles + fplu_border(n));
        // algorithmic bytes: the trailing block read and written, row and
        // column k written
        RLG_PROF_W("fplu_step", s, (2.0 * (double)wdt * (double)wdt + 2.0 * (double)n) * sizeof(T));
        hipLaunchKernelGGL(k_fplu_step<T>, dim3(grid), dim3(FTH), 0, s, a, (long)lda, (int)n, (int)k,
                           (int)tiles, w);
        RLG_LAUNCHED();
    }
    return RLG_OK;
}

int fplu_result(hipStream_t s, size_t n, void *work, size_t elem, int64_t *prow, int64_t *pcol,
                int64_t *steps) {
    *steps = 0;
    if (n == 0) return RLG_OK;
    // the step arrays and the step count sit at the same offsets for both
    // element types once carved with the right size
    int64_t *dprow, *dpcol, *dsteps;
    if (elem == 8) {
        const auto w = fplu_carve<double>(work, n);
        dprow = w.prow; dpcol = w.pcol; dsteps = w.steps;
    } else {
        const auto w = fplu_carve<float>(work, n);
        dprow = w.prow; dpcol = w.pcol; dsteps = w.steps;
    }
    RLG_HIP(hipMemcpyAsync(prow, dprow, n * 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipMemcpyAsync(pcol, dpcol, n * 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipMemcpyAsync(steps, dsteps, 8, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

// x[r][perm[i]] = y[r][i] for nrhs vectors (the `&q * x` of lu.rs:463).
template <class T>
__global__ void k_scatter_vec(int n, const int64_t *__restrict__ perm, const T *__restrict__ y, long ldy,
                              T *__restrict__ x, long ldx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (i < n) x[(long)r * ldx + perm[i]] = y[(long)r * ldy + i];
}

template <class T>
int scatter_vec(hipStream_t s, size_t n, size_t nrhs, const int64_t *perm, const T *y, size_t ldy, T *x,
                size_t ldx) {
    if (n == 0 || nrhs == 0) return RLG_OK;
    hipLaunchKernelGGL(k_scatter_vec<T>, dim3(ceil_div(n, 256), (unsigned)nrhs), dim3(256), 0, s, (int)n, perm,
                       y, (long)ldy, x, (long)ldx);
    RLG_LAUNCHED();
    return RLG_OK;
}

template int fplu_factor<double>(hipStream_t, size_t, double *, size_t, void *);
template int fplu_factor<float>(hipStream_t, size_t, float *, size_t, void *);
template int scatter_vec<double>(hipStream_t, size_t, size_t, const int64_t *, const double *, size_t, double *,
                                 size_t);
template int scatter_vec<float>(hipStream_t, size_t, size_t, const int64_t *, const float *, size_t, float *,
                                size_t);

}  // namespace rlg

// ---------------------------------------------------------------- C ABI

struct rlg_fplu {
    int elem = 8;                 // 8: f64 factor, 4: f32 factor
    size_t n = 0, lda = 0;
    size_t steps = 0;             // elimination steps taken (n unless a zero pivot ended them)
    void *a = nullptr;            // packed factor on the device
    rlg::PoolBuf own;             // backing store when the handle owns `a`
    rlg::PoolBuf rp_d, q_d, rq_d; // int64 on the device: raw row map, stored q, raw column map
    std::vector<int64_t> p, q;    // the stored permutations (lu.rs:425-429)
    std::vector<double> diag;     // U's diagonal (rank, is_invertible, det, singularity)
    mutable rlg::StreamUse use;
};

namespace rlg {
namespace {

// The raw maps from the step pivots: raw[k] = the original index now at
// position k after the swaps of steps 0..steps-1 (PermutationMatrix::
// swap_rows, permutation_matrix.rs:132-134); the stored permutation is its
// inverse (lu.rs:425-429).
void raw_from_steps(size_t n, size_t steps, const int64_t *piv, std::vector<int64_t> &raw) {
    raw.resize(n);
    for (size_t i = 0; i < n; ++i) raw[i] = (int64_t)i;
    for (size_t k = 0; k < steps; ++k) std::swap(raw[k], raw[(size_t)piv[k]]);
}

std::vector<int64_t> inverse_of(const std::vector<int64_t> &raw) {
    std::vector<int64_t> inv(raw.size());
    for (size_t i = 0; i < raw.size(); ++i) inv[(size_t)raw[i]] = (int64_t)i;
    return inv;
}

template <class T>
int fplu_factor_handle(hipStream_t s, rlg_fplu *h) {
    const size_t n = h->n;
    PoolBuf work;
    RLG_TRY(work.alloc(fplu_work_bytes(n, sizeof(T)), s));
    RLG_TRY(fplu_factor<T>(s, n, static_cast<T *>(h->a), h->lda, work.p));
    std::vector<int64_t> prow(n ? n : 1), pcol(n ? n : 1), rp, rq;
    int64_t steps = 0;
    RLG_TRY(fplu_result(s, n, work.p, sizeof(T), prow.data(), pcol.data(), &steps));
    h->steps = (size_t)steps;
    raw_from_steps(n, h->steps, prow.data(), rp);
    raw_from_steps(n, h->steps, pcol.data(), rq);
    h->p = inverse_of(rp);
    h->q = inverse_of(rq);
    std::vector<T> dg(n);
    if (n)
        RLG_HIP(hipMemcpy2DAsync(dg.data(), sizeof(T), h->a, (h->lda + 1) * sizeof(T), sizeof(T), n,
                                 hipMemcpyDeviceToHost, s));
    RLG_TRY(h->rp_d.alloc(n * 8, s));
    RLG_TRY(h->q_d.alloc(n * 8, s));
    RLG_TRY(h->rq_d.alloc(n * 8, s));
    if (n) {
        RLG_HIP(hipMemcpyAsync(h->rp_d.p, rp.data(), n * 8, hipMemcpyHostToDevice, s));
        RLG_HIP(hipMemcpyAsync(h->q_d.p, h->q.data(), n * 8, hipMemcpyHostToDevice, s));
        RLG_HIP(hipMemcpyAsync(h->rq_d.p, rq.data(), n * 8, hipMemcpyHostToDevice, s));
    }
    RLG_HIP(hipStreamSynchronize(s));
    h->diag.assign(dg.begin(), dg.end());
    return RLG_OK;
}

template <class T>
int fplu_decompose_host(size_t rows, size_t cols, const T *a, size_t lda, rlg_fplu **out) {
    if (!out) { set_error("null output pointer"); return RLG_E_INVALID_ARG; }
    *out = nullptr;
    if (rows != cols) {
        set_error("Matrix must be square for LU decomposition (%zu x %zu).", rows, cols);
        return RLG_E_SHAPE;
    }
    if (lda < cols || (!a && rows)) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    auto *h = new rlg_fplu();
    h->elem = sizeof(T);
    h->n = rows;
    h->lda = rows;
    int st = h->own.alloc(rows * rows * sizeof(T), s);
    if (st == RLG_OK) {
        h->a = h->own.p;
        if (rows)
            if (hipMemcpy2DAsync(h->a, rows * sizeof(T), a, lda * sizeof(T), rows * sizeof(T), rows,
                                 hipMemcpyHostToDevice, s) != hipSuccess) {
                set_error("upload failed");
                st = RLG_E_DEVICE;
            }
    }
    if (st == RLG_OK) st = fplu_factor_handle<T>(s, h);
    if (st != RLG_OK) { delete h; return st; }
    *out = h;
    return RLG_OK;
}

// T::epsilon, the absolute test of back_substitution (src/matrix/mod.rs:332-335).
template <class T> double eps_of();
template <> double eps_of<double>() { return 2.220446049250313e-16; }
template <> double eps_of<float>() { return 1.1920928955078125e-07; }

// FullPivLu::epsilon (lu.rs:596-598): |lu[0][0]| * T::epsilon, or T::epsilon
// for an empty matrix; computed in T.
template <class T>
double fplu_eps(const rlg_fplu *h) {
    const T e = (T)eps_of<T>();
    const T a00 = h->n ? (T)std::fabs(h->diag[0]) : (T)1;
    return (double)(a00 * e);
}

// The solve on device-resident vectors: y = P b gathered by the raw row map,
// the two substitutions, then the scatter through `perm` (the stored q for
// the reference's result, the raw column map for the true solution).
template <class T>
int fplu_solve_dev(hipStream_t s, const rlg_fplu *h, size_t nrhs, const T *b, size_t ldb, T *x, size_t ldx,
                   bool true_solution) {
    const size_t n = h->n;
    // back_substitution (src/matrix/mod.rs:330-336): a diagonal entry of U
    // below T::epsilon in magnitude fails the solve
    for (size_t i = 0; i < n; ++i)
        if ((T)std::fabs(h->diag[i]) < (T)eps_of<T>()) {
            set_error("Lower triangular matrix is singular to working precision.");
            return RLG_E_DIV_BY_ZERO;
        }
    if (nrhs == 0 || n == 0) return RLG_OK;
    PoolBuf y, tmp;
    RLG_TRY(y.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(tmp.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(lu_solve<T>(s, n, static_cast<const T *>(h->a), h->lda, h->rp_d.as<int64_t>(), nrhs, b, ldb,
                        y.as<T>(), n, tmp.as<T>()));
    RLG_TRY(scatter_vec<T>(s, n, nrhs, true_solution ? h->rq_d.as<int64_t>() : h->q_d.as<int64_t>(), y.as<T>(),
                           n, x, ldx));
    h->use.add(s);
    return RLG_OK;
}

template <class T>
int fplu_solve_host(const rlg_fplu *h, size_t b_len, size_t nrhs, const T *b, size_t ldb, T *x, size_t ldx,
                    bool true_solution) {
    if (!h || h->elem != (int)sizeof(T)) { set_error("factor handle of another precision"); return RLG_E_INVALID_ARG; }
    const size_t n = h->n;
    if (b_len != n) {
        set_error("Right-hand side vector must have compatible size (%zu vs %zu).", b_len, n);
        return RLG_E_SHAPE;
    }
    if (nrhs && n && (!b || !x || ldb < n || ldx < n)) { set_error("bad right-hand side"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s = d->stream;
    if (nrhs == 0 || n == 0) return fplu_solve_dev<T>(s, h, 0, b, ldb, x, ldx, true_solution);
    PoolBuf db, dx;
    RLG_TRY(db.alloc(nrhs * n * sizeof(T), s));
    RLG_TRY(dx.alloc(nrhs * n * sizeof(T), s));
    RLG_HIP(hipMemcpy2DAsync(db.p, n * sizeof(T), b, ldb * sizeof(T), n * sizeof(T), nrhs, hipMemcpyHostToDevice,
                             s));
    RLG_TRY(fplu_solve_dev<T>(s, h, nrhs, db.as<T>(), n, dx.as<T>(), n, true_solution));
    RLG_HIP(hipMemcpy2DAsync(x, ldx * sizeof(T), dx.p, n * sizeof(T), n * sizeof(T), nrhs, hipMemcpyDeviceToHost,
                             s));
    RLG_HIP(hipStreamSynchronize(s));
    return RLG_OK;
}

template <class T>
int fplu_unpack(const rlg_fplu *h, T *l, T *u) {
    if (!h || h->elem != (int)sizeof(T)) { set_error("factor handle of another precision"); return RLG_E_INVALID_ARG; }
    const size_t n = h->n;
    if (n == 0) return RLG_OK;
    if (!l || !u) { set_error("null output pointer"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    std::vector<T> lu(n * n);
    RLG_HIP(hipMemcpy2DAsync(lu.data(), n * sizeof(T), h->a, h->lda * sizeof(T), n * sizeof(T), n,
                             hipMemcpyDeviceToHost, d->stream));
    RLG_HIP(hipStreamSynchronize(d->stream));
    // unit_lower_triangular_part (lu.rs:644-663), nullify_lower_triangular_part
    // (src/internal_utils.rs:5-12)
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            const T v = lu[i * n + j];
            l[i * n + j] = j < i ? v : (j == i ? T(1) : T(0));
            u[i * n + j] = j < i ? T(0) : v;
        }
    return RLG_OK;
}

}  // namespace
}  // namespace rlg

using namespace rlg;

extern "C" {

int rlg_fplu_decompose_f64(size_t rows, size_t cols, const double *a, size_t lda, rlg_fplu **out) {
    return fplu_decompose_host<double>(rows, cols, a, lda, out);
}

int rlg_fplu_decompose_f32(size_t rows, size_t cols, const float *a, size_t lda, rlg_fplu **out) {
    return fplu_decompose_host<float>(rows, cols, a, lda, out);
}

int rlg_fplu_decompose_f64_dev(size_t n, double *a, size_t lda, void *stream, rlg_fplu **out) {
    if (!out) { set_error("null output pointer"); return RLG_E_INVALID_ARG; }
    *out = nullptr;
    if (lda < n || (!a && n)) { set_error("bad operand"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    auto *h = new rlg_fplu();
    h->elem = 8;
    h->n = n;
    h->lda = lda;
    h->a = a;
    int st = fplu_factor_handle<double>(pick_stream(d, stream), h);
    if (st != RLG_OK) { delete h; return st; }
    *out = h;
    return RLG_OK;
}

int rlg_fplu_solve_f64(const rlg_fplu *lu, size_t b_len, size_t nrhs, const double *b, size_t ldb, double *x,
                       size_t ldx) {
    return fplu_solve_host<double>(lu, b_len, nrhs, b, ldb, x, ldx, false);
}

int rlg_fplu_solve_f32(const rlg_fplu *lu, size_t b_len, size_t nrhs, const float *b, size_t ldb, float *x,
                       size_t ldx) {
    return fplu_solve_host<float>(lu, b_len, nrhs, b, ldb, x, ldx, false);
}

int rlg_fplu_solve_true_f64(const rlg_fplu *lu, size_t b_len, size_t nrhs, const double *b, size_t ldb,
                            double *x, size_t ldx) {
    return fplu_solve_host<double>(lu, b_len, nrhs, b, ldb, x, ldx, true);
}

int rlg_fplu_solve_f64_dev(const rlg_fplu *lu, size_t nrhs, const double *b, size_t ldb, double *x, size_t ldx,
                           int true_solution, void *stream) {
    if (!lu || lu->elem != 8) { set_error("not an f64 factor"); return RLG_E_INVALID_ARG; }
    if (nrhs && lu->n && (!b || !x || ldb < lu->n || ldx < lu->n)) {
        set_error("bad right-hand side");
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    return fplu_solve_dev<double>(pick_stream(d, stream), lu, nrhs, b, ldb, x, ldx, true_solution != 0);
}

int rlg_fplu_size(const rlg_fplu *lu, size_t *n) {
    if (!lu || !n) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    *n = lu->n;
    return RLG_OK;
}

int rlg_fplu_perms(const rlg_fplu *lu, size_t *p, size_t *q) {
    if (!lu) { set_error("null handle"); return RLG_E_INVALID_ARG; }
    for (size_t i = 0; i < lu->n; ++i) {
        if (p) p[i] = (size_t)lu->p[i];
        if (q) q[i] = (size_t)lu->q[i];
    }
    return RLG_OK;
}

int rlg_fplu_steps(const rlg_fplu *lu, size_t *steps) {
    if (!lu || !steps) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    *steps = lu->steps;
    return RLG_OK;
}

int rlg_fplu_unpack_f64(const rlg_fplu *lu, double *l, double *u) { return fplu_unpack<double>(lu, l, u); }
int rlg_fplu_unpack_f32(const rlg_fplu *lu, float *l, float *u) { return fplu_unpack<float>(lu, l, u); }

int rlg_fplu_rank(const rlg_fplu *lu, size_t *rank) {
    if (!lu || !rank) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    // lu.rs:544-557: leading diagonal entries above epsilon
    const double eps = lu->elem == 8 ? fplu_eps<double>(lu) : fplu_eps<float>(lu);
    size_t r = 0;
    for (size_t i = 0; i < lu->n; ++i) {
        if (std::fabs(lu->diag[i]) > eps) ++r;
        else break;
    }
    *rank = r;
    return RLG_OK;
}

int rlg_fplu_is_invertible(const rlg_fplu *lu, int *yes) {
    if (!lu || !yes) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    // lu.rs:582-594: the last diagonal entry above epsilon; true when empty
    const double eps = lu->elem == 8 ? fplu_eps<double>(lu) : fplu_eps<float>(lu);
    *yes = lu->n == 0 ? 1 : (std::fabs(lu->diag[lu->n - 1]) > eps ? 1 : 0);
    return RLG_OK;
}

int rlg_fplu_det_f64(const rlg_fplu *lu, double *det) {
    if (!lu || !det || lu->elem != 8) { set_error("null argument or not an f64 factor"); return RLG_E_INVALID_ARG; }
    // lu.rs:505-517: the diagonal folded from one, times the parities of p and q
    double u_det = 1.0;
    for (size_t i = 0; i < lu->n; ++i) u_det = u_det * lu->diag[i];
    auto parity = [](std::vector<int64_t> perm) {
        // permute_by_swap (permutation_matrix.rs:479-509)
        bool even = true;
        for (size_t i = 0; i < perm.size(); ++i) {
            size_t target = (size_t)perm[i];
            while (i != target) {
                const size_t next = (size_t)perm[target];
                even = !even;
                perm[target] = (int64_t)target;
                target = next;
            }
            perm[i] = (int64_t)i;
        }
        return even ? 1.0 : -1.0;
    };
    *det = parity(lu->p) * u_det * parity(lu->q);
    return RLG_OK;
}

int rlg_fplu_inverse_f64(const rlg_fplu *lu, double *inv, size_t ld) {
    if (!lu || lu->elem != 8) { set_error("not an f64 factor"); return RLG_E_INVALID_ARG; }
    const size_t n = lu->n;
    int ok = 0;
    RLG_TRY(rlg_fplu_is_invertible(lu, &ok));
    if (!ok) {  // lu.rs:477-482
        set_error("Non-invertible matrix found while attempting inversion");
        return RLG_E_DIV_BY_ZERO;
    }
    if (n == 0) return RLG_OK;
    if (!inv || ld < n) { set_error("bad output"); return RLG_E_INVALID_ARG; }
    // lu.rs:484-494: one solve per unit vector, stored as column i
    std::vector<double> e(n * n, 0.0), cols(n * n);
    for (size_t i = 0; i < n; ++i) e[i * n + i] = 1.0;
    RLG_TRY(rlg_fplu_solve_f64(lu, n, n, e.data(), n, cols.data(), n));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) inv[j * ld + i] = cols[i * n + j];
    return RLG_OK;
}

void rlg_fplu_free(rlg_fplu *lu) {
    if (!lu) return;
    const auto streams = lu->use.all();
    lu->own.release_after(streams);
    lu->rp_d.release_after(streams);
    lu->q_d.release_after(streams);
    lu->rq_d.release_after(streams);
    delete lu;
}

}  // extern "C"
