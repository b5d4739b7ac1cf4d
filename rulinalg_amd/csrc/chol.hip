// Blocked Cholesky (Cholesky::decompose, src/matrix/decomposition/cholesky.rs:
// 116-169), lower triangle only, and its solve (cholesky.rs:194-202).
//
// Outer blocks of CONB = 512 columns end in one SYRK-shaped MFMA GEMM of
// depth 512 on the lower tiles of the trailing matrix (512 measured faster
// than 128 or 256 at order 16384, and wider blocks no faster).  Inside an outer
// block, 64-column blocks are processed left to right:
//   1. one launch (k_chol_block) factors the diagonal block with the
//      reference's left-looking "gaxpy" recurrence: a[k][j] -= dot(row k,
//      row j) over the block's previous columns in utils::dot's eight-lane
//      order, the |d| < eps and d < 0 checks, and division by sqrt(d);
//      in the same launch the rows below are solved against it with the
//      same recurrence (a triangular solve X L^T = B, row by row), one
//      column behind;
//   2. a depth-64 GEMM updates the remaining columns of the outer block.
// A matrix of order <= 64 is therefore factored bit-identically to the
// reference.  Only entries on or below the diagonal are read; entries above
// it may be overwritten with intermediate values (the reference's unpack
// zeroes them, cholesky.rs:237-246).
#include <utility>

#include "kernels.hpp"
#include "wave.hpp"

namespace rlg {

constexpr int CNB = 64;
constexpr int CONB = 512;

template <class T> __device__ __forceinline__ T ceps();
template <> __device__ __forceinline__ double ceps<double>() { return 2.220446049250313e-16; }
template <> __device__ __forceinline__ float ceps<float>() { return 1.1920929e-7f; }

// utils::dot (src/utils.rs:20-51) of x[0..LEN) held in registers and
// l[0..LEN) in LDS, LEN known at compile time: partial sums p0..p7 over
// chunks of eight, combined as s + p0 + p4 + p1 + p5 + p2 + p6 + p3 + p7,
// then the tail in order.
template <int LEN, class T>
__device__ __forceinline__ T dot8_reg(const T (&x)[CNB], const T *l) {
#pragma clang fp contract(off)
    constexpr int M8 = LEN / 8 * 8;
    T p[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) p[q] = T(0);
#pragma unroll
    for (int off = 0; off < M8; off += 8) {
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] = p[q] + x[off + q] * l[off + q];
    }
    T s = T(0);
    s = s + p[0] + p[4];
    s = s + p[1] + p[5];
    s = s + p[2] + p[6];
    s = s + p[3] + p[7];
#pragma unroll
    for (int o = M8; o < LEN; ++o) s = s + x[o] * l[o];
    return s;
}

// One 64-column block column j0 .. j0+jb: the diagonal block and the rows
// below it, in one launch of two-wave workgroups.
//
// Wave 0 factors the jb x jb diagonal block with the reference's
// left-looking recurrence (cholesky.rs:135-163): a[k][j] -= dot(a[k][0..j],
// a[j][0..j]) for k >= j in utils::dot order, the |d| < eps and d < 0
// checks, then division by sqrt(d).  Every workgroup does this redundantly
// (the same arithmetic, so the same values).  The block is written back in
// place, so it may be written only once every workgroup has read it: the
// workgroups count themselves in `arrive` once their loads are done, and the
// last to arrive writes the block and the divisors (a workgroup may start
// long after the others when other streams' kernels hold the CUs, and then
// must still read the unfactored block).  Workgroup 0 writes the failure
// index, which no workgroup reads.  Wave 1 solves the workgroup's 64
// rows below the block, x L11^T = b, with the same recurrence,
// x[c] = (b[c] - dot(x[0..c], L11[c][0..c])) / divisor[c], one column
// behind wave 0: column c needs only row c of L11 and divisor c, both final
// once wave 0 has finished column c.  Each thread keeps its row in
// registers (the column loop is expanded at compile time, so every index is
// a constant) and reads the broadcast row of L11 from LDS.  The rows below
// are read and written whole (coalesced) and transposed through LDS.
template <class T>
__global__ __launch_bounds__(128) void k_chol_block(T *__restrict__ A, long lda, int j0, int jb,
                                                    int r0, int n, T *__restrict__ divisor,
                                                    int64_t *__restrict__ info,
                                                    unsigned *__restrict__ arrive) {
#pragma clang fp contract(off)
    __shared__ T a[CNB][CNB + 1];
    __shared__ T X[64][CNB + 1];
    __shared__ T sdiv[CNB];
    __shared__ int last;  // this workgroup is the last to have read the block
    const int tid = threadIdx.x, t = tid & 63;
    const bool dw = tid < 64;  // wave 0: diagonal block; wave 1: rows below
    const int rb = r0 + (int)blockIdx.x * 64;
    const int nr = n - rb >= 64 ? 64 : (n - rb > 0 ? n - rb : 0);
    {   // all loads in flight before their LDS stores
        T v[64];
        if (dw) {
#pragma unroll
            for (int r = 0; r < CNB; ++r) v[r] = (r < jb && t <= r) ? A[(long)(j0 + r) * lda + j0 + t] : T(0);
#pragma unroll
            for (int r = 0; r < CNB; ++r)
                if (r < jb && t <= r) a[r][t] = v[r];
        } else {
#pragma unroll
            for (int r = 0; r < 64; ++r) v[r] = (r < nr && t < jb) ? A[(long)(rb + r) * lda + j0 + t] : T(0);
#pragma unroll
            for (int r = 0; r < 64; ++r) X[r][t] = v[r];
        }
    }
    __syncthreads();
    // every load of this workgroup has returned (its values are in LDS)
    if (tid == 0)
        last = __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
               gridDim.x - 1;
    T x[CNB];  // wave 0: row t of the diagonal block; wave 1: row rb + t
#pragma unroll
    for (int c = 0; c < CNB; ++c) x[c] = dw ? (c <= t ? a[t][c] : T(0)) : X[t][c];
    bool failed = false;
    int jend = jb;  // columns whose divisor was computed
    static_for<CNB + 1>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        // wave 1: column j-1 of its rows (row j-1 of L11 and divisor j-1 are final)
        if constexpr (j > 0) {
            if (!dw && j - 1 < jb && !failed)
                x[j - 1] = (x[j - 1] - (j > 1 ? dot8_reg<j - 1>(x, &a[j - 1][0]) : T(0))) / sdiv[j - 1];
        }
        if constexpr (j < CNB) {
            if (j < jb && !failed) {  // the reference stops at the first failure
                if (dw && j > 0 && t >= j && t < jb) {
                    x[j] = x[j] - dot8_reg<j>(x, &a[j][0]);
                    a[t][j] = x[j];
                }
                __syncthreads();
                const T diag = a[j][j];
                if (fabs(diag) < ceps<T>() || diag < T(0)) {
                    if (tid == 0 && blockIdx.x == 0 && *info == (int64_t)n) *info = j0 + j;
                    failed = true;
                    jend = j + 1;
                }
                const T div = sqrt(diag);
                // the divided diagonal stays in thread j's register until the
                // end: every thread of both waves reads the undivided a[j][j]
                // between the two barriers, so nothing may overwrite it there
                if (dw && t >= j && t < jb) {
                    x[j] = x[j] / div;
                    if (t != j) a[t][j] = x[j];
                }
                // kept in LDS: a global store here would make the barrier wait for it
                if (tid == 0) sdiv[j] = div;
                __syncthreads();
            }
        }
    });
    if (dw) {
#pragma unroll
        for (int c = 0; c < CNB; ++c)
            if (c == t && c < jend) a[t][t] = x[c];
        if (last) {  // set before the loop's first barrier (jb >= 1)
            if (t < jend) divisor[j0 + t] = sdiv[t];
#pragma unroll 8
            for (int r = 0; r < jb; ++r)
                if (t <= r) A[(long)(j0 + r) * lda + j0 + t] = a[r][t];
        }
    } else {
#pragma unroll
        for (int c = 0; c < CNB; ++c) X[t][c] = x[c];
    }
    __syncthreads();
    if (!dw && t < jb) {
#pragma unroll 8
        for (int r = 0; r < nr; ++r) A[(long)(rb + r) * lda + j0 + t] = X[r][t];
    }
}

// Workspace: the divisors, then one arrival counter per 64-column block.
static size_t chol_div_bytes(size_t n, size_t elem) { return ((n + 1) * elem + 255) & ~(size_t)255; }
size_t chol_work_bytes(size_t n, size_t elem) {
    return chol_div_bytes(n, elem) + ((ceil_div(n, CNB) * sizeof(unsigned) + 255) & ~(size_t)255);
}

// LDS the look-ahead's update workgroups reserve beyond their own 66 KB
// (FP64): two of them no longer fit on a CU's 160 KB, one of them and a
// block-column workgroup (67 KB) do, so the block columns find a place on
// every CU while the update runs.
constexpr size_t CHOL_UPDATE_PAD = 24 << 10;

struct CholEvents {
    hipEvent_t ready = nullptr, done = nullptr;
    ~CholEvents() {
        if (ready) (void)hipEventDestroy(ready);
        if (done) (void)hipEventDestroy(done);
    }
};

// Look-ahead: after outer block k, the trailing update is split.  The next
// outer block's columns (all rows below it) are updated first, on the
// caller's stream, which then factors that block; the rest of the update
// runs meanwhile on the auxiliary stream.  The two touch disjoint columns,
// and the caller's stream waits for the rest before its next update of the
// same region, so every entry receives the same updates in the same order as
// without look-ahead; the split follows the 128-wide tile grid of the
// update (outer blocks are four tiles wide), so each entry is computed with
// the same arithmetic as well.
template <class T>
static int chol_impl(hipStream_t s, size_t n, T *a, size_t lda, int64_t *info, void *work,
                     int lowp_prec) {
    T *divisor = static_cast<T *>(work);
    unsigned *arrive = reinterpret_cast<unsigned *>(static_cast<char *>(work) + chol_div_bytes(n, sizeof(T)));
    RLG_TRY(fill_i64(info, 1, (int64_t)n, s));
    RLG_TRY(fill_u32(arrive, ceil_div(n, CNB), 0u, s));
    const bool native = lowp_prec != RLG_PREC_F16 && lowp_prec != RLG_PREC_BF16;
    AuxStream aux;
    hipStream_t s2 = nullptr;
    CholEvents ev;
    bool pending = false;  // an update of columns beyond the next block is running on s2
    // from order 8192 on: below it factorizations mostly run several at a
    // time (the serving loop), and those already share the device
    if (native && n >= 8192) {
        RLG_TRY(aux.create());
        s2 = aux.s;
        RLG_HIP(hipEventCreateWithFlags(&ev.ready, hipEventDisableTiming));
        RLG_HIP(hipEventCreateWithFlags(&ev.done, hipEventDisableTiming));
    }
    for (size_t k0 = 0; k0 < n; k0 += CONB) {
        const size_t kb = n - k0 < (size_t)CONB ? n - k0 : CONB;
        for (size_t j0 = k0; j0 < k0 + kb; j0 += CNB) {
            const size_t jb = k0 + kb - j0 < (size_t)CNB ? k0 + kb - j0 : CNB;
            const size_t r0 = j0 + jb;
            {
                RLG_PROF("chol_block", s);
                const size_t below = n - r0;  // rows under the diagonal block
                hipLaunchKernelGGL(k_chol_block<T>, dim3(below ? ceil_div(below, 64) : 1), dim3(128), 0,
                                   s, a, (long)lda, (int)j0, (int)jb, (int)r0, (int)n, divisor, info,
                                   arrive + j0 / CNB);
                RLG_LAUNCHED();
            }
            if (r0 < k0 + kb) {
                // remaining columns of the outer block, all rows below the diagonal
                RLG_TRY(gemm_t<T>(s, n - r0, k0 + kb - r0, jb, T(-1), a + r0 * lda + j0, lda,
                                  a + r0 * lda + j0, lda, T(1), a + r0 * lda + r0, lda, true, false));
            }
        }
        const size_t c0 = k0 + kb;
        if (c0 >= n) break;
        if (pending) {  // the previous step's update of the columns beyond this block
            RLG_HIP(hipStreamWaitEvent(s, ev.done, 0));
            pending = false;
        }
        const T *lk = a + c0 * lda + k0;  // this block's columns of L, rows from c0
        if (!native) {
            RLG_TRY(gemm_lowp(s, lowp_prec, n - c0, n - c0, kb, -1.f, reinterpret_cast<const float *>(lk),
                              lda, reinterpret_cast<const float *>(lk), lda, 1.f,
                              reinterpret_cast<float *>(a + c0 * lda + c0), lda, true, true));
            continue;
        }
        const size_t kb2 = n - c0 < (size_t)CONB ? n - c0 : CONB;
        if (!s2 || c0 + kb2 >= n) {
            RLG_TRY(gemm_t<T>(s, n - c0, n - c0, kb, T(-1), lk, lda, lk, lda, T(1), a + c0 * lda + c0, lda,
                              true, true));
            continue;
        }
        // the next block's columns first (rows c0 .. n-1, lower part of its diagonal block)
        RLG_TRY(gemm_t<T>(s, n - c0, kb2, kb, T(-1), lk, lda, lk, lda, T(1), a + c0 * lda + c0, lda, true,
                          true));
        // the rest on the auxiliary stream, once this block's columns are final
        const size_t c1 = c0 + kb2;
        RLG_HIP(hipEventRecord(ev.ready, s));
        RLG_HIP(hipStreamWaitEvent(s2, ev.ready, 0));
        RLG_TRY(gemm_t_padded<T>(s2, n - c1, n - c1, kb, T(-1), a + c1 * lda + k0, lda,
                                 a + c1 * lda + k0, lda, T(1), a + c1 * lda + c1, lda, true, true,
                                 CHOL_UPDATE_PAD));
        RLG_HIP(hipEventRecord(ev.done, s2));
        pending = true;
    }
    if (pending) RLG_HIP(hipStreamWaitEvent(s, ev.done, 0));
    return RLG_OK;
}

template <class T>
int chol_factor(hipStream_t s, size_t n, T *a, size_t lda, int64_t *info, void *work) {
    return chol_impl<T>(s, n, a, lda, info, work, RLG_PREC_F64);
}
template int chol_factor<double>(hipStream_t, size_t, double *, size_t, int64_t *, void *);
template int chol_factor<float>(hipStream_t, size_t, float *, size_t, int64_t *, void *);

int chol_factor_lowp(hipStream_t s, int prec, size_t n, float *a, size_t lda, int64_t *info,
                     void *work) {
    return chol_impl<float>(s, n, a, lda, info, work, prec);
}

// Cholesky::solve (cholesky.rs:194-202): forward substitution with L, then
// transpose_back_substitution with L^T.  b is copied to tmp first, so x may
// alias b; the first solve writes x, the second reads x and writes tmp.
template <class T>
int chol_solve(hipStream_t s, size_t n, const T *l, size_t lda, size_t nrhs, const T *b, size_t ldb,
               T *x, size_t ldx, T *tmp) {
    if (n == 0 || nrhs == 0) return RLG_OK;
    RLG_HIP(hipMemcpy2DAsync(tmp, n * sizeof(T), b, ldb * sizeof(T), n * sizeof(T), nrhs,
                             hipMemcpyDeviceToDevice, s));
    RLG_TRY(trsv_lower<T>(s, n, l, lda, false, nrhs, tmp, n, x, ldx));
    RLG_TRY(trsv_lower_t<T>(s, n, l, lda, nrhs, x, ldx, tmp, n));
    RLG_HIP(hipMemcpy2DAsync(x, ldx * sizeof(T), tmp, n * sizeof(T), n * sizeof(T), nrhs,
                             hipMemcpyDeviceToDevice, s));
    return RLG_OK;
}
template int chol_solve<double>(hipStream_t, size_t, const double *, size_t, size_t, const double *,
                                size_t, double *, size_t, double *);
template int chol_solve<float>(hipStream_t, size_t, const float *, size_t, size_t, const float *,
                               size_t, float *, size_t, float *);

}  // namespace rlg
