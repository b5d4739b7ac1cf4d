// Triangular solves on nrhs right-hand sides (vector r at in + r * ldi),
// one launch per step of TSB x 64 rows.
//
// A step owns the rows [s0, s1) of one "super-block" (TSB sub-blocks of 64
// rows, taken in solve order).  Every workgroup of the launch first solves
// the super-block redundantly: it stages the super-block's part of the
// triangle in LDS and one wave per right-hand side runs the substitution
// with the vector held in registers (lane i = row i of a sub-block) and the
// solved entry broadcast lane to lane, so no workgroup ever waits for
// another.  Then each workgroup removes the contribution of the solved rows
// from its share of the rows still to be solved (a GEMV slice: rows below
// the super-block for forward substitution, rows above it for backward).
// Workgroup 0 writes the solved rows to `out`.
//
// `in` holds the right-hand sides and is used as the running residual: a
// step reads the rows of its super-block (final since the previous step)
// and updates only rows outside it, so reads and writes of one launch never
// touch the same element.  The solution goes to `out`, a different buffer.
//
// Within a sub-block the order of operations is the column sweep of the
// reference's transpose_back_substitution (src/matrix/decomposition/
// cholesky.rs:329-361): divide x[c] by the diagonal, then subtract
// m[i][c] * x[c] (separately rounded) from the remaining rows.
// forward_substitution, lu_forward_substitution and back_substitution
// (src/matrix/mod.rs:318-398, lu.rs:624-642) form row dot products instead;
// the results agree to rounding.
#include "kernels.hpp"
#include "wave.hpp"

namespace rlg {

constexpr int SB = 64;      // sub-block rows (the lanes of one wave)
constexpr int TSB = 2;      // sub-blocks per step
constexpr int TMAXR = 16;   // right-hand sides per launch
constexpr int RPW = 64;     // rows updated per workgroup (row-access forms)
constexpr int CPW = 64;     // entries updated per workgroup (transposed form)

enum TriMode { LOWER_UNIT = 0, LOWER = 1, UPPER = 2, LOWER_T = 3 };

// x / d on the solve's serial chain: FP64 divides (the reference divides);
// FP32 (the refinement's correction solves) multiplies by a reciprocal that
// every lane computed up front.
__device__ __forceinline__ double diag_div(double x, double d, double) { return x / d; }
__device__ __forceinline__ float diag_div(float x, float, float rd) { return x * rd; }

template <int MODE> __device__ __forceinline__ constexpr bool forward() {
    return MODE == LOWER_UNIT || MODE == LOWER;
}
template <int MODE> __device__ __forceinline__ constexpr bool stored_upper() { return MODE == UPPER; }

// One 64 x 64 tile of M (rows rb.., cols cb.., rs x cs valid), 16 values per
// thread: tile_load issues the loads, tile_store puts them in LDS.  `diag`
// keeps only the stored triangle and zeroes the rest, so entries of the
// other triangle are never read as numbers.  With `vec` (M and lda allow
// 16-byte loads) a full tile is read with 16-byte loads, a quarter (FP32)
// or half (FP64) of the load instructions: with the update slice also in
// flight, scalar loads would exceed what a wave can have outstanding.
// Staged rows are padded to 16-byte multiples (TP) so that a lane can read
// its row back with 16-byte LDS reads.
constexpr int TPER = SB * SB / 256;
template <class T> constexpr int TVL = 16 / sizeof(T);  // elements per 16 bytes
template <class T> constexpr int TP = TVL<T>;           // row padding
template <class T, int MODE>
__device__ __forceinline__ void tile_load(T (&v)[TPER], const T *__restrict__ M, long lda, int rb,
                                          int cb, int rs, int cs, bool diag, bool vec) {
    constexpr int VL = TVL<T>;
    typedef T vT __attribute__((ext_vector_type(VL)));
    if (vec && rs == SB && cs == SB) {
#pragma unroll
        for (int u = 0; u < TPER / VL; ++u) {
            const int e = u * 256 + threadIdx.x;  // vector index
            const int i = e / (SB / VL), j0 = (e % (SB / VL)) * VL;
            const vT x = *reinterpret_cast<const vT *>(M + (long)(rb + i) * lda + cb + j0);
#pragma unroll
            for (int t = 0; t < VL; ++t) {
                const int j = j0 + t;
                const bool ok = !diag || (stored_upper<MODE>() ? j >= i : j <= i);
                v[u * VL + t] = ok ? x[t] : T(0);
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < TPER; ++q) {
            const int e = q * 256 + threadIdx.x;
            const int i = e >> 6, j = e & 63;
            bool ok = i < rs && j < cs;
            if (diag) ok = ok && (stored_upper<MODE>() ? j >= i : j <= i);
            v[q] = ok ? M[(long)(rb + i) * lda + cb + j] : T(0);
        }
    }
}
template <class T>
__device__ __forceinline__ void tile_store(T (*tl)[SB + TP<T>], const T (&v)[TPER], bool vec_full) {
    constexpr int VL = TVL<T>;
    if (vec_full) {
#pragma unroll
        for (int u = 0; u < TPER / VL; ++u) {
            const int e = u * 256 + threadIdx.x;
            const int i = e / (SB / VL), j0 = (e % (SB / VL)) * VL;
#pragma unroll
            for (int t = 0; t < VL; ++t) tl[i][j0 + t] = v[u * VL + t];
        }
    } else {
#pragma unroll
        for (int q = 0; q < TPER; ++q) {
            const int e = q * 256 + threadIdx.x;
            tl[e >> 6][e & 63] = v[q];
        }
    }
}

// Row `lane` of a staged tile into registers (r[c] = t[lane][c]), or for the
// transposed form its column `lane` (r[c] = t[c][lane]): the serial chain of
// the substitution then reads registers, not LDS.
template <class T, int MODE>
__device__ __forceinline__ void tile_row(const T (*t)[SB + TP<T>], int lane, T (&r)[SB]) {
    constexpr int VL = TVL<T>;
    typedef T vT __attribute__((ext_vector_type(VL)));
    if (MODE == LOWER_T) {
#pragma unroll
        for (int c = 0; c < SB; ++c) r[c] = t[c][lane];
    } else {
#pragma unroll
        for (int c = 0; c < SB; c += VL) {
            const vT x = *reinterpret_cast<const vT *>(&t[lane][c]);
#pragma unroll
            for (int q = 0; q < VL; ++q) r[c + q] = x[q];
        }
    }
}

// Right-hand sides whose update-phase rows of W are loaded up front.
constexpr int RPRE = 2;

template <class T, int MODE>
__global__ __launch_bounds__(256) void k_trsv_step(const T *__restrict__ M, long lda, int n, int q0,
                                                   int nq, int nrhs, T *__restrict__ W, long ldw,
                                                   T *__restrict__ X, long ldx, int u0, int u1,
                                                   int vec) {
    // tiles: [0] diagonal of sub-block 0, [1] diagonal of sub-block 1,
    // [2] the off-diagonal tile inside the stored triangle
    __shared__ __attribute__((aligned(16))) T tl[3][SB][SB + TP<T>];
    __shared__ T xs[TMAXR][TSB * SB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s0 = q0 * SB, s1 = min(n, (q0 + nq) * SB), S = s1 - s0;
    const int bs0 = min(SB, S), bs1 = S - bs0;  // rows of sub-blocks 0 and 1
    constexpr bool fwd = forward<MODE>();

    // ---- every read that does not depend on the super-block's solution is
    // issued first, so a launch waits for memory once, not once per phase:
    // the update slice and, for the first right-hand sides, its rows of W;
    // the three tiles; this wave's first right-hand side on the super-block.
    constexpr int NIT = RPW / 16, NC = TSB * SB / 16;  // row forms
    constexpr int CW = TSB * SB / 4;                   // transposed form
    const int ls = lane & 15, grp = lane >> 4;
    const int rbase = u0 + (int)blockIdx.x * RPW + wave * (RPW / 4) + grp;
    const int jt = u0 + (int)blockIdx.x * CPW + lane;
    T mv[MODE == LOWER_T ? CW : NIT * NC];
    T wpre[RPRE][MODE == LOWER_T ? 1 : NIT];
    if constexpr (MODE == LOWER_T) {
        const bool act = jt < u1;
#pragma unroll
        for (int q = 0; q < CW; ++q) {
            const int c = wave * CW + q;
            mv[q] = (act && c < S) ? M[(long)(s0 + c) * lda + jt] : T(0);
        }
#pragma unroll
        for (int r = 0; r < RPRE; ++r)
            wpre[r][0] = (wave == 0 && act && r < nrhs) ? W[(long)r * ldw + jt] : T(0);
    } else {
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int j = rbase + it * 4;
            const T *mrow = M + (long)j * lda + s0;
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const int c = ls + 16 * k;
                mv[it * NC + k] = (j < u1 && c < S) ? mrow[c] : T(0);
            }
        }
#pragma unroll
        for (int r = 0; r < RPRE; ++r)
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int j = rbase + it * 4;
                wpre[r][it] = (ls == 0 && j < u1 && r < nrhs) ? W[(long)r * ldw + j] : T(0);
            }
    }
    {
        T v0[TPER], v1[TPER], v2[TPER];
        const bool vf0 = vec && bs0 == SB, vf1 = vec && bs1 == SB;  // full tiles
        tile_load<T, MODE>(v0, M, lda, s0, s0, bs0, bs0, true, vec);
        if (nq > 1) {
            tile_load<T, MODE>(v1, M, lda, s0 + SB, s0 + SB, bs1, bs1, true, vec);
            if (stored_upper<MODE>())
                tile_load<T, MODE>(v2, M, lda, s0, s0 + SB, bs0, bs1, false, vec);
            else
                tile_load<T, MODE>(v2, M, lda, s0 + SB, s0, bs1, bs0, false, vec);
        }
        // this wave's first right-hand side on the two sub-blocks
        const bool r0ok = wave < nrhs;
        const T xw0 = (r0ok && lane < bs0) ? W[(long)wave * ldw + s0 + lane] : T(0);
        const T xw1 = (r0ok && nq > 1 && lane < bs1) ? W[(long)wave * ldw + s0 + SB + lane] : T(0);
        tile_store(tl[0], v0, vf0);
        if (nq > 1) {
            tile_store(tl[1], v1, vf1);
            tile_store(tl[2], v2, vf0 && vf1);
        }
        __syncthreads();

        // ---- 1. the super-block, redundantly in every workgroup
        for (int r = wave; r < nrhs; r += 4) {
#pragma clang fp contract(off)
            T xq0 = T(0), xq1 = T(0);
            for (int qi = 0; qi < nq; ++qi) {
                const int q = fwd ? qi : nq - 1 - qi;
                const int bs = q == 0 ? bs0 : bs1;
                T x = r == wave ? (q == 0 ? xw0 : xw1)
                                : (lane < bs ? W[(long)r * ldw + s0 + q * SB + lane] : T(0));
                // The loops below run all 64 steps, unrolled, so the LDS reads of
                // the tile are issued ahead of the serial chain.  Steps beyond the
                // block size change nothing: the staged tiles are zero outside the
                // block and x is zero on lanes beyond it, so they subtract exact
                // zeros.
                if (qi == 1) {  // contribution of the sub-block solved first
                    const T xp = fwd ? xq0 : xq1;
                    T m[SB];
                    tile_row<T, MODE>(tl[2], lane, m);
#pragma unroll
                    for (int c = 0; c < SB; ++c) {
                        const T xc = lane_get(xp, c);
                        x = x - m[c] * xc;
                    }
                }
                const T(*d)[SB + TP<T>] = tl[q];
                T dr[SB];  // this lane's row (column for the transposed form) of d
                tile_row<T, MODE>(d, lane, dr);
                const T dl = lane < bs ? d[lane][lane] : T(1);
                const T rd = T(1) / dl;  // used by the FP32 form only
                if (fwd) {
#pragma unroll
                    for (int c = 0; c < SB; ++c) {
                        if (MODE == LOWER && lane == c) x = diag_div(x, dl, rd);
                        const T xc = lane_get(x, c);
                        if (lane > c) x = x - dr[c] * xc;
                    }
                } else {
#pragma unroll
                    for (int cc = 0; cc < SB; ++cc) {
                        const int c = SB - 1 - cc;
                        if (lane == c) x = diag_div(x, dl, rd);
                        const T xc = lane_get(x, c);
                        if (lane < c) x = x - dr[c] * xc;
                    }
                }
                if (lane >= bs) x = T(0);
                if (q == 0) xq0 = x; else xq1 = x;
            }
            xs[r][lane] = xq0;
            xs[r][SB + lane] = xq1;
        }
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = tid; e < nrhs * S; e += 256) {
            const int r = e / S, c = e - r * S;
            X[(long)r * ldx + s0 + c] = xs[r][c];
        }

    // ---- 2. remove the solved rows from this workgroup's share of [u0, u1)
    if constexpr (MODE == LOWER_T) {
        // W[r][j] -= sum_c L[s0 + c][j] * x[c]: a workgroup owns 64 entries j
        // (one per lane: coalesced rows of L), wave w sums c in [32w, 32w + 32)
        const bool act = jt < u1;
        // partial sums meet in LDS (the staged tiles are no longer read)
        T(*red)[TMAXR][SB] = reinterpret_cast<T(*)[TMAXR][SB]>(&tl[0][0][0]);
        for (int r = 0; r < nrhs; ++r) {
            T a = T(0);
#pragma unroll
            for (int q = 0; q < CW; ++q) a += mv[q] * xs[r][wave * CW + q];
            red[wave][r][lane] = a;
        }
        __syncthreads();
        if (wave == 0 && act) {
            for (int r = 0; r < nrhs; ++r) {
                const T a = (red[0][r][lane] + red[1][r][lane]) + (red[2][r][lane] + red[3][r][lane]);
                const T w0 = r < RPRE ? wpre[r < RPRE ? r : 0][0] : W[(long)r * ldw + jt];
                W[(long)r * ldw + jt] = w0 - a;
            }
        }
    } else {
        // W[r][j] -= sum_c M[j][s0 + c] * x[c]; 16 lanes per row, 4 rows per
        // wave at a time, lane ls holds columns ls, ls + 16, ... (coalesced)
        for (int r = 0; r < nrhs; ++r) {  // wave-uniform
            T v[NIT];
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                T a = T(0);
#pragma unroll
                for (int k = 0; k < NC; ++k) a += mv[it * NC + k] * xs[r][ls + 16 * k];
                v[it] = row16_sum(a);  // the 16 lanes of a row are one DPP row
            }
            if (ls == 0) {
                T w[NIT];
#pragma unroll
                for (int it = 0; it < NIT; ++it) {
                    const int j = rbase + it * 4;
                    w[it] = r < RPRE ? wpre[r < RPRE ? r : 0][it] : (j < u1 ? W[(long)r * ldw + j] : T(0));
                }
#pragma unroll
                for (int it = 0; it < NIT; ++it) {
                    const int j = rbase + it * 4;
                    if (j < u1) W[(long)r * ldw + j] = w[it] - v[it];
                }
            }
        }
    }
}

template <class T, int MODE>
static int trsv_run(hipStream_t s, size_t n, const T *m, size_t lda, size_t nrhs, T *w, size_t ldw,
                    T *x, size_t ldx) {
    if (n == 0 || nrhs == 0) return RLG_OK;
    if (w == x) {
        set_error("triangular solve: residual and output buffers must differ");
        return RLG_E_INVALID_ARG;
    }
    constexpr bool fwd = MODE == LOWER_UNIT || MODE == LOWER;
    // 16-byte tile loads when rows of the triangle start 16-byte aligned
    const bool vec = lda % TVL<T> == 0 && (reinterpret_cast<uintptr_t>(m) & 15) == 0;
    const size_t nb = (n + SB - 1) / SB;
    const size_t steps = (nb + TSB - 1) / TSB;
    for (size_t r0 = 0; r0 < nrhs; r0 += TMAXR) {
        const size_t nr = nrhs - r0 < (size_t)TMAXR ? nrhs - r0 : TMAXR;
        for (size_t k = 0; k < steps; ++k) {
            size_t q0, nq;
            if (fwd) {
                q0 = k * TSB;
                nq = nb - q0 < (size_t)TSB ? nb - q0 : TSB;
            } else {
                const size_t qe = nb - k * TSB;
                q0 = qe > (size_t)TSB ? qe - TSB : 0;
                nq = qe - q0;
            }
            const size_t s0 = q0 * SB, s1 = (q0 + nq) * SB < n ? (q0 + nq) * SB : n;
            const size_t u0 = fwd ? s1 : 0, u1 = fwd ? n : s0;
            const size_t per = MODE == LOWER_T ? CPW : RPW;
            const unsigned grid = u1 > u0 ? ceil_div(u1 - u0, per) : 1u;
            // algorithmic bytes: the super-block triangle, the update slice,
            // the right-hand-side rows read and written
            const double bytes =
                sizeof(T) * ((double)(s1 - s0) * (s1 - s0 + 1) / 2 + (double)(u1 - u0) * (s1 - s0) +
                             (double)nr * (2.0 * (u1 - u0) + 2.0 * (s1 - s0)));
            RLG_PROF_W("trsv", s, bytes);
            hipLaunchKernelGGL((k_trsv_step<T, MODE>), dim3(grid), dim3(256), 0, s, m, (long)lda,
                               (int)n, (int)q0, (int)nq, (int)nr, w + r0 * ldw, (long)ldw,
                               x + r0 * ldx, (long)ldx, (int)u0, (int)u1, vec ? 1 : 0);
            RLG_LAUNCHED();
        }
    }
    return RLG_OK;
}

template <class T>
int trsv_lower(hipStream_t s, size_t n, const T *l, size_t lda, bool unit, size_t nrhs, T *in,
               size_t ldi, T *out, size_t ldo) {
    return unit ? trsv_run<T, LOWER_UNIT>(s, n, l, lda, nrhs, in, ldi, out, ldo)
                : trsv_run<T, LOWER>(s, n, l, lda, nrhs, in, ldi, out, ldo);
}

template <class T>
int trsv_upper(hipStream_t s, size_t n, const T *u, size_t lda, size_t nrhs, T *in, size_t ldi,
               T *out, size_t ldo) {
    return trsv_run<T, UPPER>(s, n, u, lda, nrhs, in, ldi, out, ldo);
}

template <class T>
int trsv_lower_t(hipStream_t s, size_t n, const T *l, size_t lda, size_t nrhs, T *in, size_t ldi,
                 T *out, size_t ldo) {
    return trsv_run<T, LOWER_T>(s, n, l, lda, nrhs, in, ldi, out, ldo);
}

// The singularity test of back_substitution and forward_substitution
// (src/matrix/mod.rs:331-335, 375-378): a diagonal entry below epsilon in
// magnitude.  Every flagged lane stores the same value with a vector store.
template <class T>
__global__ void k_diag_tiny(const T *__restrict__ M, long lda, long n, T eps, uint32_t *flag) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && std::fabs(M[i * lda + i]) < eps) *flag = 1u;
}

template <class T>
int diag_tiny(hipStream_t s, size_t n, const T *m, size_t lda, T eps, uint32_t *flag) {
    if (n == 0) return RLG_OK;
    hipLaunchKernelGGL(k_diag_tiny<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, m, (long)lda,
                       (long)n, eps, flag);
    RLG_LAUNCHED();
    return RLG_OK;
}
template int diag_tiny<double>(hipStream_t, size_t, const double *, size_t, double, uint32_t *);
template int diag_tiny<float>(hipStream_t, size_t, const float *, size_t, float, uint32_t *);

#define RLG_TRSV_INST(T)                                                                           \
    template int trsv_lower<T>(hipStream_t, size_t, const T *, size_t, bool, size_t, T *, size_t, \
                               T *, size_t);                                                       \
    template int trsv_upper<T>(hipStream_t, size_t, const T *, size_t, size_t, T *, size_t, T *,   \
                               size_t);                                                            \
    template int trsv_lower_t<T>(hipStream_t, size_t, const T *, size_t, size_t, T *, size_t, T *, \
                                 size_t);
RLG_TRSV_INST(double)
RLG_TRSV_INST(float)

}  // namespace rlg
