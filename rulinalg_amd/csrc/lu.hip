// Right-looking blocked LU with partial pivoting (PartialPivLu::decompose,
// src/matrix/decomposition/lu.rs:163-194) and its solve (lu.rs:231-244).
//
// Blocking: outer blocks of ONB = 512 columns drive the trailing Schur
// update as one MFMA GEMM of depth 512 (faster than 256 at order 16384 in
// both FP64 and FP32: less C traffic per flop); inside an outer block, inner panels
// of PNB = 64 columns are factored by a kernel launched cooperatively
// (hipLaunchCooperativeKernel: all of its workgroups resident together) and
// applied to the rest of the outer block by a 64-row triangular solve and a
// depth-64 GEMM.
//
// The panel kernel keeps the whole (n - k) x 64 panel resident in LDS,
// spread over G workgroups of PR rows each.  Every column costs one
// exchange: each workgroup publishes its local pivot candidate (|value|,
// row, and the full 64-wide row) and the content of the current diagonal
// row as tagged words; every workgroup reads the G candidates, reduces them
// to the same winner, performs its part of the row swap, and eliminates its
// rows.  The pivot rule is the reference's: the first row of maximal
// magnitude at or below the diagonal (strict `>` scan, lu.rs:170-178); the
// multiplier is a division and the update a separately rounded multiply and
// subtract (lu.rs:603-616), so a matrix no wider than one panel is factored
// bit-identically to the reference.
#include <climits>
#include <type_traits>

#include "kernels.hpp"
#include "mfma.hpp"
#include "wave.hpp"

namespace rlg {

constexpr int PNB = 64;    // inner panel width
// Panel rows per workgroup (LDS-resident): 256 (133 KB of LDS in FP64, one
// workgroup per CU, so every workgroup of a panel is resident at once).
constexpr int PR = 256;
template <class T> struct PanelRows { static constexpr int value = PR; };
template <> struct PanelRows<float> { static constexpr int value = 2 * PR; };  // same 133 KB
constexpr int ONB = 512;   // outer block: depth of the trailing GEMM
constexpr int SWB = 256;   // swaps per k_laswp launch (its workgroup size)
constexpr int PTHREADS = 256;
// At most this many workgroups per panel: one per CU of the device and never
// more than the exchange area holds.
constexpr int GMAX = 256;

template <class T> __device__ __forceinline__ T eps_of();
template <> __device__ __forceinline__ double eps_of<double>() { return 2.220446049250313e-16; }
template <> __device__ __forceinline__ float eps_of<float>() { return 1.1920929e-7f; }

// Exchange words of the panel workgroups.  A 64-bit word carries 32 bits of
// payload in its low half and, in its high half, the tag of the column it
// belongs to: global column + 1, 0 = never written in this factorization.
// Words are written and read with agent-scope atomics, and a reader polls
// until every word it needs carries the current tag, so a value is never
// used before it is complete and no counter, fence or barrier is involved.
// A slot holds a header (candidate row, candidate |value|) and a 64-wide row;
// a value takes one word in FP32 and two (low, high half) in FP64.
typedef unsigned long long xword;
constexpr int XHDR = 4;  // header words: row, |v| (1 or 2 words), spare
template <class T> struct XSlot {
    static constexpr int VW = sizeof(T) / 4;                      // words per value
    static constexpr int words = (XHDR + PNB * VW + 15) & ~15;    // 128-byte multiple
};
// Slots per column parity: one per workgroup, then the diagonal-row slot.
constexpr int XPAR = GMAX + 1;

// Workspace layout (bytes): exchange slots, inverses of the unit-lower
// diagonal blocks of the current outer block, timeout flag.
struct LuWork {
    xword *xch;        // [2][XPAR][XSlot<T>::words]
    void *linv;        // [ONB / PNB][PNB][PNB] T
    int *timeout;      // exchange timeout flag
};

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }
static size_t xch_words(size_t elem) {
    return 2 * (size_t)XPAR * (elem == 8 ? XSlot<double>::words : XSlot<float>::words);
}

size_t lu_work_bytes(size_t n, size_t elem) {
    (void)n;
    return align_up(xch_words(elem) * sizeof(xword)) + align_up((size_t)ONB * PNB * elem) +
           align_up(4);
}

static LuWork carve(void *work, size_t n, size_t elem) {
    (void)n;
    char *p = static_cast<char *>(work);
    LuWork w;
    w.xch = reinterpret_cast<xword *>(p); p += align_up(xch_words(elem) * sizeof(xword));
    w.linv = p; p += align_up((size_t)ONB * PNB * elem);
    w.timeout = reinterpret_cast<int *>(p);
    return w;
}

int *lu_timeout_flag(void *work, size_t n, size_t elem) { return carve(work, n, elem).timeout; }

void perm_from_ipiv(size_t n, const int64_t *ipiv, int64_t *raw) {
    for (size_t i = 0; i < n; ++i) raw[i] = (int64_t)i;
    for (size_t k = 0; k < n; ++k) {
        const size_t p = (size_t)ipiv[k];
        if (p != k) { int64_t t = raw[k]; raw[k] = raw[p]; raw[p] = t; }
    }
}

// ---------------------------------------------------------------- panel

__device__ __forceinline__ unsigned xtag(xword w) { return (unsigned)(w >> 32); }
__device__ __forceinline__ void xput(xword *p, unsigned tag, unsigned v) {
    __hip_atomic_store(p, ((xword)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ xword xget(const xword *p) {
    return __hip_atomic_load(const_cast<xword *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xput_val(xword *p, unsigned tag, float v) {
    xput(p, tag, __float_as_uint(v));
}
__device__ __forceinline__ void xput_val(xword *p, unsigned tag, double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    xput(p, tag, (unsigned)b);
    xput(p + 1, tag, (unsigned)(b >> 32));
}
// The value at p; ok = every word of it carries `tag`.
template <class T> __device__ __forceinline__ T xget_val(const xword *p, unsigned tag, bool &ok);
template <> __device__ __forceinline__ float xget_val<float>(const xword *p, unsigned tag, bool &ok) {
    const xword a = xget(p);
    ok = xtag(a) == tag;
    return __uint_as_float((unsigned)a);
}
template <> __device__ __forceinline__ double xget_val<double>(const xword *p, unsigned tag, bool &ok) {
    const xword a = xget(p), b = xget(p + 1);
    ok = xtag(a) == tag && xtag(b) == tag;
    return __longlong_as_double((long long)(((b & 0xffffffffull) << 32) | (a & 0xffffffffull)));
}

// One round of a bounded wave-wide wait.  Returns false once the timeout
// flag is up (checked every 256 rounds) or after 2^22 rounds, when it raises
// the flag itself: an exchange that cannot complete then ends the kernel,
// and the host reports the flag, instead of hanging.  The result is uniform
// over the wave.
__device__ __forceinline__ bool xwait(unsigned &spins, int *timeout) {
    if ((++spins & 255u) == 0) {
        const int up = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (up) return false;
        if (spins >= (1u << 22)) {
            if ((threadIdx.x & 63) == 0)
                __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    __builtin_amdgcn_s_sleep(1);
    return true;
}

// Factor the panel A[k:k+m, k:k+w] (w <= PNB) in place.
//
// Workgroup g holds rows [g PRT, (g+1) PRT) of the panel in LDS; thread t
// owns the tile rows t + h PTHREADS and does every local step on them.  The
// elimination is lazy by one column: pivot j is applied to column j + 1 at
// once, and to the columns beyond during the next column's exchange, while
// the workgroup waits for the other candidates anyway.  Column j:
//   A. every thread finds the first maximum |a[r][j]| over its rows r >= j
//      (column j is up to date); each wave copies its candidate row, and the
//      wave holding row j that row, to LDS; barrier;
//   B. wave 0 picks the workgroup's candidate and brings the copied rows up
//      to date with pivot j - 1 (the arithmetic the owners apply in place),
//      then publishes the candidate (row index, |value|, the row) and, if it
//      holds row j, row j;
//   C/D1. pivot j - 1 is applied to columns > j of every row; with other
//      workgroups (G > 1, FP32) wave 0 meanwhile reads all G candidates, picks the
//      same winner as every other workgroup and fetches the winning row
//      unless it is its own (and, if it holds the winning row but not row j,
//      row j), while waves 1-3 update all rows, wave 0's included; alone,
//      every thread updates its own rows; barrier;
//   D2. wave 0 swaps the rows it holds and puts the pivot row in LDS;
//      barrier;
//   E. every thread turns column j of its rows below j into multipliers and
//      applies pivot j to column j + 1.
// The pivot indices stay in LDS until the end: a global store before a
// workgroup barrier would make the barrier wait for its completion.

// Tile rows are padded to a 16-byte multiple: 16-byte aligned rows let the
// row sweeps use 16-byte LDS accesses.
template <class T> struct TilePad { static constexpr int value = 16 / sizeof(T); };
template <class T> constexpr int TW = PNB + TilePad<T>::value;

// a[r][c] = a[r][c] - mult * pm[c] for the columns c > j from chunk C0 on,
// as one straight sweep: the row part is loaded, then stored.  The first
// 8-column chunk may hold columns <= j.  With EXACT they keep their value,
// by a select on the computed value, as in the reference, which never
// touches them (x - mult * 0 would turn a -0 into +0, and a non-finite
// multiplier would reach them).  Without EXACT they take x - mult * 0 (pm is
// zero there), which differs from x only in the sign of a zero: the mixed
// factorization, whose factors only precondition the FP64 refinement, uses
// that form, which needs fewer registers.  Every column is stored
// unconditionally: a per-column test around the store gets compiled into a
// conditional store, one LDS round trip per column.  Columns w..63 of the
// tile and of pm hold zeros and stay zero.
template <class T, int C0, bool EXACT>
__device__ __forceinline__ void elim_row(T (*__restrict__ tile)[TW<T>], const T *__restrict__ pm,
                                         int r, int j, T mult) {
#pragma clang fp contract(off)
    T x[PNB - C0];
#pragma unroll
    for (int c = C0; c < PNB; ++c) x[c - C0] = tile[r][c];
#pragma unroll
    for (int c = C0; c < PNB; ++c) {
        const T v = x[c - C0] - mult * pm[c];
        tile[r][c] = (!EXACT || c >= C0 + 8 || c > j) ? v : x[c - C0];
    }
}

// elim_row from the 8-column chunk that holds column j + 1.
template <class T, bool EXACT>
__device__ __forceinline__ void elim_row_from(T (*__restrict__ tile)[TW<T>], const T *__restrict__ pm,
                                              int r, int j, T mult) {
    switch ((j + 1) >> 3) {
        case 0: elim_row<T, 0, EXACT>(tile, pm, r, j, mult); break;
        case 1: elim_row<T, 8, EXACT>(tile, pm, r, j, mult); break;
        case 2: elim_row<T, 16, EXACT>(tile, pm, r, j, mult); break;
        case 3: elim_row<T, 24, EXACT>(tile, pm, r, j, mult); break;
        case 4: elim_row<T, 32, EXACT>(tile, pm, r, j, mult); break;
        case 5: elim_row<T, 40, EXACT>(tile, pm, r, j, mult); break;
        case 6: elim_row<T, 48, EXACT>(tile, pm, r, j, mult); break;
        case 7: elim_row<T, 56, EXACT>(tile, pm, r, j, mult); break;
        default: break;  // j = 63: nothing right of column j
    }
}

template <class T, bool EXACT>
__global__ __launch_bounds__(PTHREADS) void k_lu_panel(T *__restrict__ A, long lda, int k, int m,
                                                       int w, int64_t *__restrict__ ipiv,
                                                       int64_t *__restrict__ info, int n,
                                                       xword *__restrict__ xch, int *timeout) {
    constexpr int PRT = PanelRows<T>::value;
    constexpr int RPT = PRT / PTHREADS;  // panel rows per thread
    static_assert(RPT * PTHREADS == PRT, "whole rows per thread");
    constexpr int NW = PTHREADS / 64;
    constexpr int VW = XSlot<T>::VW, XS = XSlot<T>::words;
    __shared__ __attribute__((aligned(16))) T tile[PRT][TW<T>];
    __shared__ __attribute__((aligned(16))) T prow[PNB];      // pivot row j
    // pivot row j with zeros at columns <= j + 1, by column parity (its
    // pending update is applied during column j + 1's step)
    __shared__ __attribute__((aligned(16))) T pmask[2][PNB];
    __shared__ T wrow[NW][PNB];                                // wave candidates' rows
    __shared__ T jrow[PNB];                                    // row j
    __shared__ T red_v[NW];
    __shared__ int red_i[NW];
    __shared__ int s_piv[PNB];

    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int r0 = g * PRT;
    const int rows = min(PRT, m - r0);
    const int lane = tid & 63, wave = tid >> 6;

    {   // all loads in flight before the LDS stores (PRT x PNB / PTHREADS per thread)
        constexpr int PER = PRT * PNB / PTHREADS;
        T v[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * PTHREADS + tid, r = e / PNB, c = e % PNB;
            v[q] = (r < rows && c < w) ? A[(long)(k + r0 + r) * lda + k + c] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * PTHREADS + tid, r = e / PNB, c = e % PNB;
            if (r < rows) tile[r][c] = v[q];  // columns w..63 zero
        }
    }
    __syncthreads();

    bool live = true;  // wave 0: false once an exchange has timed out
    int small = w;     // wave 0: first column whose pivot is below epsilon
    for (int j = 0; j < w; ++j) {
        const T *pmprev = pmask[(j + 1) & 1];  // masked pivot row j - 1
        const bool own_j = j >= r0 && j < r0 + rows;

        // A. this thread's candidate, the wave's, and the rows wave 0 needs
        T bv = T(0);
        int bi = INT_MAX;
#pragma unroll
        for (int h = 0; h < RPT; ++h) {
            const int r = tid + h * PTHREADS;
            const T v = fabs(tile[r][j]);
            const bool take = (r < rows) & (r0 + r >= j) & better(v, r0 + r, bv, bi);
            bv = take ? v : bv;
            bi = take ? r0 + r : bi;
        }
        {
            const int wl = wave_first_max_lane(bv, bi);
            bv = wl >= 0 ? lane_get(bv, wl) : T(0);
            bi = wl >= 0 ? lane_get(bi, wl) : INT_MAX;
        }
        {   // the wave's candidate row and row j, brought up to date with pivot
            // j - 1 right of column j (the arithmetic step C applies to the
            // rows in place); columns <= j are final and stay as they are
#pragma clang fp contract(off)
            if (bi != INT_MAX) {
                const T x = tile[bi - r0][lane];
                const T v = x - tile[bi - r0][j > 0 ? j - 1 : 0] * pmprev[lane];
                wrow[wave][lane] = (j > 0 && lane > j) ? v : x;
            }
            if (own_j && (((j - r0) % PTHREADS) >> 6) == wave) {
                const T x = tile[j - r0][lane];
                const T v = x - tile[j - r0][j > 0 ? j - 1 : 0] * pmprev[lane];
                jrow[lane] = (j > 0 && lane > j) ? v : x;
            }
        }
        if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; }
        __syncthreads();

        // B. the workgroup's candidate row and row j, published
        T pv = T(0), oj = T(0);
        if (wave == 0) {
            int bw = 0;
#pragma unroll
            for (int q = 1; q < NW; ++q) {
                const bool take = better(red_v[q], red_i[q], bv, bi);
                bv = take ? red_v[q] : bv;
                bi = take ? red_i[q] : bi;
                bw = take ? q : bw;
            }
            if (bi != INT_MAX) pv = wrow[bw][lane];
            if (own_j) oj = jrow[lane];
            if (G > 1 && live) {
                const unsigned tag = (unsigned)(k + j + 1);
                xword *base = xch + (size_t)(j & 1) * XPAR * XS;
                xword *mine = base + (size_t)g * XS, *diag = base + (size_t)GMAX * XS;
                if (lane < w) {
                    if (bi != INT_MAX) xput_val(mine + XHDR + lane * VW, tag, pv);
                    if (own_j) xput_val(diag + XHDR + lane * VW, tag, oj);
                }
                if (lane == 0) {
                    xput(mine, tag, (unsigned)bi);
                    xput_val(mine + 1, tag, bv);
                }
            }
        }

        // D1 / C.  With other workgroups to hear from, wave 0 waits for the
        // exchange (D1) while waves 1-3 apply pivot j - 1 to every row,
        // wave 0's included (C); alone, each wave updates its own rows.
        // (FP32 only: with two rows per thread the hand-over spreads evenly;
        // with one, a wave of 1-3 would do twice its work, measured slower;
        // FP64 updates its rows first and then exchanges, after the barrier)
        const bool w0_exchange = RPT > 1 && G > 1;  // uniform
        int p = bi;
        T pabs = bv;
        auto exchange = [&]() {
            if (live) {
                const unsigned tag = (unsigned)(k + j + 1);
                const xword *base = xch + (size_t)(j & 1) * XPAR * XS;
                const xword *diag = base + (size_t)GMAX * XS;
                // every workgroup's candidate, 64 slots per round
                T v = T(0);
                int i = INT_MAX, gw = 0;
                for (int q0 = 0; q0 < G && live; q0 += 64) {
                    const int q = q0 + lane;
                    const xword *h = base + (size_t)q * XS;
                    T qv = T(0);
                    int qi = INT_MAX;
                    for (unsigned spins = 0;;) {
                        bool ok = true;
                        if (q < G) {
                            const xword hr = xget(h);
                            qv = xget_val<T>(h + 1, tag, ok);
                            qi = (int)(unsigned)hr;
                            ok = ok && xtag(hr) == tag;
                        }
                        if (__all(ok)) break;
                        if (!xwait(spins, timeout)) { live = false; break; }
                    }
                    const bool take = (q < G) & better(qv, qi, v, i);
                    v = take ? qv : v;
                    i = take ? qi : i;
                    gw = take ? q : gw;
                }
                {
                    const int wl = wave_first_max_lane(v, i);
                    if (wl >= 0) {
                        v = lane_get(v, wl);
                        i = lane_get(i, wl);
                        gw = lane_get(gw, wl);
                    } else if (live) {  // no candidate anywhere: cannot happen; report it
                        if (lane == 0)
                            __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        live = false;
                    }
                }
                if (live) {
                    p = i;
                    pabs = v;
                    const bool need_row = gw != g;
                    const bool need_j = p != j && p >= r0 && p < r0 + rows && !own_j;
                    if (need_row || need_j) {
                        for (unsigned spins = 0;;) {
                            bool ok = true, okj = true;
                            if (lane < w) {
                                if (need_row) {
                                    const T x = xget_val<T>(base + (size_t)gw * XS + XHDR + lane * VW, tag, ok);
                                    if (ok) pv = x;
                                }
                                if (need_j) {
                                    const T x = xget_val<T>(diag + XHDR + lane * VW, tag, okj);
                                    if (okj) oj = x;
                                }
                            }
                            if (__all(ok && okj)) break;
                            if (!xwait(spins, timeout)) { live = false; break; }
                        }
                    }
                }
                // after a timeout the result is discarded: keep indices in range
                if (!live) { p = bi; pabs = bv; }
            }
        };
        if (wave == 0 && w0_exchange) exchange();
        if (j > 0) {
            if (!w0_exchange) {
#pragma unroll
                for (int h = 0; h < RPT; ++h) {
                    const int r = tid + h * PTHREADS;
                    if (r < rows && r0 + r >= j) elim_row_from<T, EXACT>(tile, pmprev, r, j, tile[r][j - 1]);
                }
            } else if (wave > 0) {
                // own rows, then wave 0's 64 RPT rows spread over the 192 threads
#pragma unroll
                for (int h = 0; h < RPT; ++h) {
                    const int r = tid + h * PTHREADS;
                    if (r < rows && r0 + r >= j) elim_row_from<T, EXACT>(tile, pmprev, r, j, tile[r][j - 1]);
                }
                constexpr int XTRA = (64 * RPT + PTHREADS - 64 - 1) / (PTHREADS - 64);
#pragma unroll
                for (int q = 0; q < XTRA; ++q) {
                    const int e = (tid - 64) + q * (PTHREADS - 64);
                    const int r = (e % 64) + (e / 64) * PTHREADS;
                    if (e < 64 * RPT && r < rows && r0 + r >= j) elim_row_from<T, EXACT>(tile, pmprev, r, j, tile[r][j - 1]);
                }
            }
        }
        __syncthreads();

        // D2. (FP64: the exchange,) the swap and the pivot row
        if (wave == 0) {
            if (G > 1 && !w0_exchange) exchange();
            // swap rows j and p where this workgroup holds them (lanes w..63
            // carry zeros on both sides)
            if (p != j) {
                if (own_j) tile[j - r0][lane] = pv;
                if (p >= r0 && p < r0 + rows) tile[p - r0][lane] = oj;
            }
            prow[lane] = pv;
            pmask[j & 1][lane] = lane > j + 1 ? pv : T(0);
            if (lane == 0) s_piv[j] = p;
            if (pabs < eps_of<T>() && small == w) small = j;
        }
        __syncthreads();

        // E. multipliers of column j; pivot j on column j + 1
        {
#pragma clang fp contract(off)
            const T piv = prow[j];
            const T pn = j + 1 < PNB ? prow[j + 1] : T(0);
#pragma unroll
            for (int h = 0; h < RPT; ++h) {
                const int r = tid + h * PTHREADS;
                if (r < rows && r0 + r > j) {
                    const T mult = tile[r][j] / piv;
                    tile[r][j] = mult;
                    if (j + 1 < PNB) tile[r][j + 1] = tile[r][j + 1] - mult * pn;
                }
            }
        }
    }
    __syncthreads();

    if (g == 0) {
        if (tid < w) ipiv[k + tid] = (int64_t)k + s_piv[tid];
        if (tid == 0 && small < w && *info == (int64_t)n) *info = k + small;
    }
    constexpr int PER = PRT * PNB / PTHREADS;
#pragma unroll 8
    for (int q = 0; q < PER; ++q) {
        const int e = q * PTHREADS + tid, r = e / PNB, c = e % PNB;
        if (r < rows && c < w) A[(long)(k + r0 + r) * lda + k + c] = tile[r][c];
    }
}

// ---------------------------------------------------------------- row swaps

// Net effect of the swaps ipiv[k0 .. k0+cnt) as a list of (dst, src) rows:
// after the swaps, row dst holds what row src held before.  Swap s exchanges
// rows k0+s and p_s >= k0+s, so row k0+s is final after swap s and then
// holds what row p_s held just before it.  Before swap t, a row x >= k0+t
// holds what row k0+s held before swap s, for the last s < t with p_s = x,
// and its own content if there is none.  Following that rule back is one
// downward scan of the swaps from t-1, so every thread resolves its rows on
// its own: thread i the final content of row k0+i and, when swap i is the
// last to touch a row p_i beyond the block, that row's (what row k0+i held
// before swap i).  The list (block rows first) is written with ballot
// prefix counts.
// The scans read P eight entries at a time (one LDS latency per eight, not
// per entry); the entries are then tested in scan order.
__device__ __forceinline__ int swap_chase(const int *P, int k0, int x, int t) {
    int j = t - 1;
    for (; j >= 7; j -= 8) {
        int p[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] = P[j - q];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (p[q] == x) x = k0 + (j - q);
    }
    for (; j >= 0; --j)
        if (P[j] == x) x = k0 + j;
    return x;
}
// True when no swap after t touches row x (P[t+1 .. cnt) != x).
__device__ __forceinline__ bool swap_last(const int *P, int x, int t, int cnt) {
    bool last = true;
    int j = t + 1;
    for (; j + 7 < cnt; j += 8) {
        int p[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] = P[j + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) last = last && p[q] != x;
    }
    for (; j < cnt; ++j) last = last && P[j] != x;
    return last;
}

// The list into LDS (dst[], src[]) by the 256 threads of a workgroup, in
// the order above; returns its length (at most 2 cnt).  P is scratch.
__device__ int swap_map_lds(const int64_t *__restrict__ ipiv, int k0, int cnt, int *P, int *dst,
                            int *src, int (*wcount)[SWB / 64]) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < cnt) P[t] = (int)ipiv[k0 + t];
    __syncthreads();
    int src_in = k0 + t, dst_out = -1, src_out = -1;
    if (t < cnt) {
        const int pt = P[t];
        src_in = swap_chase(P, k0, pt, t);
        if (pt >= k0 + cnt && swap_last(P, pt, t, cnt)) {
            dst_out = pt;
            src_out = swap_chase(P, k0, k0 + t, t);
        }
    }
    const bool mv_in = t < cnt && src_in != k0 + t;
    const bool mv_out = dst_out >= 0 && src_out != dst_out;
    const unsigned long long b_in = __ballot(mv_in), b_out = __ballot(mv_out);
    if (lane == 0) { wcount[0][wave] = __popcll(b_in); wcount[1][wave] = __popcll(b_out); }
    __syncthreads();
    int off_in = 0, tot_in = 0, off_out = 0, tot_out = 0;
#pragma unroll
    for (int q = 0; q < SWB / 64; ++q) {
        if (q < wave) { off_in += wcount[0][q]; off_out += wcount[1][q]; }
        tot_in += wcount[0][q];
        tot_out += wcount[1][q];
    }
    if (mv_in) {
        const int at = off_in + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b_in >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned)b_in, 0));
        dst[at] = k0 + t;
        src[at] = src_in;
    }
    if (mv_out) {
        const int at = tot_in + off_out +
                       (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b_out >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)b_out, 0));
        dst[at] = dst_out;
        src[at] = src_out;
    }
    __syncthreads();
    return tot_in + tot_out;
}

// Applies the swaps ipiv[k0 .. k0+cnt) (cnt <= SWB) to columns [c0, c1) and
// [c2, c3).  Every workgroup derives the row map itself (a few microseconds
// of its 256 threads, where a separate launch would cost more), then owns
// LCOLS columns: it reads every source row of the map into LDS, then writes
// the destinations, so the in-place permutation needs no ordering.
constexpr int LCOLS = 32;
template <class T>
__global__ __launch_bounds__(SWB) void k_laswp(T *__restrict__ A, long lda, const int64_t *__restrict__ ipiv,
                                               int k0, int cnt_sw, int c0, int c1, int c2, int c3) {
    __shared__ T stage[2 * SWB][LCOLS];
    __shared__ int dst[2 * SWB], src[2 * SWB];
    __shared__ int P[SWB];
    __shared__ int wcount[2][SWB / 64];
    const int cnt = swap_map_lds(ipiv, k0, cnt_sw, P, dst, src, wcount);
    if (cnt == 0) return;
    // column block -> absolute columns
    const int w1 = c1 - c0;
    const int cb = blockIdx.x * LCOLS;
    const int cl = threadIdx.x % LCOLS, rg = threadIdx.x / LCOLS;  // 8 row groups
    const int lc = cb + cl;
    const int col = lc < w1 ? c0 + lc : c2 + (lc - w1);
    const bool okc = lc < w1 + (c3 - c2);
    // eight loads in flight per thread before their LDS stores
    constexpr int RG = SWB / LCOLS;
    for (int i0 = rg; i0 < cnt; i0 += 8 * RG) {
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = i0 + q * RG;
            v[q] = (okc && i < cnt) ? A[(long)src[i] * lda + col] : T(0);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = i0 + q * RG;
            if (i < cnt) stage[i][cl] = v[q];
        }
    }
    __syncthreads();
    for (int i = rg; i < cnt; i += RG)
        if (okc) A[(long)dst[i] * lda + col] = stage[i][cl];
}

// ---------------------------------------------------------------- TRSM

// The block-row solve B := L^-1 B of a factored panel, with L the unit-lower
// jb x jb block A[j0:j0+jb, j0:j0+jb] (jb <= 64) and B = A[j0:j0+jb, c0:c1)
// (the reference applies the same elimination to these columns row by row,
// lu.rs:603-616), is done as a product with L^-1: the inverse is formed once
// per panel, and every 64-column slab of B is then one 64-deep matrix-core
// product instead of a 64-step dependent sweep.  L has a unit diagonal and,
// under partial pivoting, entries of magnitude at most 1.

// Row-major 64 x 64 inv = L^-1, zero outside jb x jb, by halves: with L = [L11 0; L21 L22] in 32 x 32 blocks,
// L^-1 = [X11 0; -X22 L21 X11, X22], X11 = L11^-1 and X22 = L22^-1.  Waves 0
// and 1 invert the two diagonal halves side by side, lane j forward-
// substituting column j of the identity with the column in registers and L
// from LDS as broadcasts; the corner is then two 32 x 32 matrix-core
// products, one 16 x 16 tile per wave.  The rows and columns past jb stage
// as zeros, so they invert to the identity and are masked on the way out.
template <class T>
__global__ __launch_bounds__(256) void k_tri_inv(const T *__restrict__ A, long lda, int j0, int jb,
                                                 T *__restrict__ inv) {
    using Acc = typename Mfma<T>::acc;
    constexpr int H = PNB / 2;
    __shared__ T L[PNB][PNB + 1];  // L, overwritten by L^-1 block by block
    __shared__ T Tm[H][H + 1];     // L21 X11
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    {
        constexpr int PER = PNB * PNB / 256;
        T v[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e >> 6, c = e & 63;
            v[q] = (r < jb && c < jb) ? A[(long)(j0 + r) * lda + j0 + c] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e >> 6, c = e & 63;
            L[r][c] = v[q];
        }
    }
    __syncthreads();
    if (wave < 2 && lane < H) {
        // X11 (wave 0) or X22 (wave 1): every read of the block precedes the
        // wave's writes to it, and the two blocks do not overlap
        const int o = wave * H, j = lane;
        T x[H];
#pragma unroll
        for (int i = 0; i < H; ++i) {
            T s = (i == j) ? T(1) : T(0);
#pragma unroll
            for (int q = 0; q < i; ++q) s -= L[o + i][o + q] * x[q];
            x[i] = s;
        }
#pragma unroll
        for (int i = 0; i < H; ++i) L[o + i][o + j] = x[i];  // zero above the diagonal
    }
    __syncthreads();
    const int tr = (wave >> 1) * 16, tc = (wave & 1) * 16;  // this wave's 16 x 16 tile
    {   // Tm = L21 X11
        Acc acc = Acc{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < H / 4; ++s)
            acc = Mfma<T>::op(L[H + tr + lr][4 * s + lk], L[4 * s + lk][tc + lr], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) Tm[tr + Mfma<T>::row(lk, r)][tc + lr] = acc[r];
    }
    __syncthreads();
    {   // X21 = -X22 Tm, over L21 (no longer read)
        Acc acc = Acc{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < H / 4; ++s)
            acc = Mfma<T>::op(L[H + tr + lr][H + 4 * s + lk], Tm[4 * s + lk][tc + lr], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) L[H + tr + Mfma<T>::row(lk, r)][tc + lr] = -acc[r];
    }
    __syncthreads();
    {
        constexpr int PER = PNB * PNB / 256;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, i = e >> 6, c = e & 63;
            inv[e] = (i < jb && c < jb && c <= i) ? L[i][c] : T(0);
        }
    }
}

// B := inv * B for one 64-column slab per workgroup, in place: the slab is
// read whole into LDS before any of it is written.  Wave w computes rows
// [16 w, 16 w + 16) of the slab as four 16 x 16 accumulators over 16 k-steps.
template <class T>
__global__ __launch_bounds__(256) void k_trsm_llu(T *__restrict__ A, long lda, int j0, int jb, int c0,
                                                  int c1, const T *__restrict__ inv) {
    using Acc = typename Mfma<T>::acc;
    __shared__ T Li[PNB][PNB + 1];  // Li[k][i] = inv[i][k]
    __shared__ T X[PNB][PNB + 1];   // X[k][c] = B[k][c]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int cb = c0 + blockIdx.x * 64;
    const int nc = min(64, c1 - cb);
    {
        constexpr int PER = PNB * PNB / 256;
        T vl[PER], vx[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e >> 6, c = e & 63;
            vl[q] = inv[e];
            vx[q] = (r < jb && c < nc) ? A[(long)(j0 + r) * lda + cb + c] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e >> 6, c = e & 63;
            Li[c][r] = vl[q];
            X[r][c] = vx[q];
        }
    }
    __syncthreads();
    Acc acc[4];
#pragma unroll
    for (int jc = 0; jc < 4; ++jc) acc[jc] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < PNB / 4; ++s) {
        const T a = Li[4 * s + lk][16 * wave + lr];
#pragma unroll
        for (int jc = 0; jc < 4; ++jc) acc[jc] = Mfma<T>::op(a, X[4 * s + lk][16 * jc + lr], acc[jc]);
    }
#pragma unroll
    for (int jc = 0; jc < 4; ++jc)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * wave + Mfma<T>::row(lk, r), col = 16 * jc + lr;
            if (row < jb && col < nc) A[(long)(j0 + row) * lda + cb + col] = acc[jc][r];
        }
}

// ---------------------------------------------------------------- driver

// Set by lu_factor_lowp for its duration: FP32 panels are then factored with
// tournament pivoting (tslu.hip) in this workspace.
static thread_local void *g_tslu_work = nullptr;
// Cleared by lu_factor_lowp for its duration: the mixed factorization's
// panels may turn a -0 left of the pivot column into +0 (see elim_row); the
// drop-in factorizations keep the reference's zeros exactly.
static thread_local bool g_panel_exact = true;

template <class T>
static const void *panel_kernel() {
    return g_panel_exact ? reinterpret_cast<const void *>(k_lu_panel<T, true>)
                         : reinterpret_cast<const void *>(k_lu_panel<T, false>);
}

template <class T>
static int panel(hipStream_t s, T *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
                 int64_t *info, size_t n, const LuWork &lw);

template <class T>
static T *linv_of(const LuWork &lw, size_t j0);

// The panel by tournament pivoting when lu_factor_lowp asked for it (the
// tournament also writes L11^-1: *linv_done), else by partial pivoting.
template <class T>
static int panel_any(hipStream_t s, T *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
                     int64_t *info, size_t n, const LuWork &lw, bool *linv_done) {
    *linv_done = false;
    if constexpr (std::is_same<T, float>::value) {
        if (g_tslu_work && m <= tslu_max_rows(n)) {
            *linv_done = true;
            return tslu_panel(s, a, lda, k, m, w, ipiv, info, n, g_tslu_work, linv_of<float>(lw, k));
        }
    }
    return panel<T>(s, a, lda, k, m, w, ipiv, info, n, lw);
}

// Largest panel height: the panel's workgroups exchange data while they run,
// so all of them must be resident at once, and their slots must fit the
// exchange area.  The bound is the device's occupancy for this kernel (one
// workgroup per CU: its LDS tile takes most of the CU's LDS) times its CUs.
template <class T>
static int panel_rows_max(size_t *out) {
    Device *d = nullptr;
    RLG_TRY(current_device(&d));
    static thread_local int per_cu[2][2] = {{-1, -1}, {-1, -1}};
    int &bpc = per_cu[sizeof(T) == 8][g_panel_exact];
    if (bpc < 0)
        RLG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, panel_kernel<T>(), PTHREADS, 0));
    size_t g = (size_t)bpc * (size_t)d->num_cus;
    if (g > (size_t)GMAX) g = GMAX;
    *out = g * PanelRows<T>::value;
    return RLG_OK;
}

template <class T>
static int panel(hipStream_t s, T *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
                 int64_t *info, size_t n, const LuWork &lw) {
    size_t mmax = 0;
    RLG_TRY(panel_rows_max<T>(&mmax));
    if (m > mmax) {
        set_error("LU panel of %zu rows exceeds the %zu rows that this device's CUs hold", m, mmax);
        return RLG_E_INVALID_ARG;
    }
    const unsigned G = ceil_div(m, PanelRows<T>::value);
    // algorithmic bytes: the m x w panel read once and written once
    RLG_PROF_W("lu_panel", s, 2.0 * (double)m * (double)w * sizeof(T));
    // A cooperative launch: the runtime starts the grid only with every
    // workgroup resident at once (and refuses a grid that cannot be), which
    // the per-column exchange between workgroups relies on.
    T *a_ = a;
    long lda_ = (long)lda;
    int k_ = (int)k, m_ = (int)m, w_ = (int)w, n_ = (int)n;
    xword *xch_ = lw.xch;
    int *to_ = lw.timeout;
    void *args[] = {&a_, &lda_, &k_, &m_, &w_, &ipiv, &info, &n_, &xch_, &to_};
    RLG_HIP(hipLaunchCooperativeKernel(panel_kernel<T>(), dim3(G), dim3(PTHREADS), args, 0, s));
    return RLG_OK;
}

template <class T>
static int laswp(hipStream_t s, T *a, size_t lda, const int64_t *ipiv, size_t k0, size_t cnt,
                 size_t c0, size_t c1, size_t c2, size_t c3) {
    const size_t ncols = (c1 - c0) + (c3 - c2);
    if (ncols == 0 || cnt == 0) return RLG_OK;
    // a longer sequence is applied SWB swaps at a time, in order: the same
    // sequence of transpositions
    for (size_t off = 0; off < cnt; off += SWB) {
        const size_t c = cnt - off < (size_t)SWB ? cnt - off : SWB;
        // algorithmic bytes: at most 2 c rows change, each read and written
        RLG_PROF_W("laswp", s, 2.0 * 2.0 * (double)c * (double)ncols * sizeof(T));
        hipLaunchKernelGGL(k_laswp<T>, dim3(ceil_div(ncols, LCOLS)), dim3(SWB), 0, s, a, (long)lda,
                           ipiv, (int)(k0 + off), (int)c, (int)c0, (int)c1, (int)c2, (int)c3);
        RLG_LAUNCHED();
    }
    return RLG_OK;
}

// The inverse of panel j0's unit-lower block: slot (j0 mod ONB) / PNB of the
// outer block's inverses.
template <class T>
static T *linv_of(const LuWork &lw, size_t j0) {
    return static_cast<T *>(lw.linv) + ((j0 % ONB) / PNB) * PNB * PNB;
}

template <class T>
static int tri_inv(hipStream_t s, const T *a, size_t lda, size_t j0, size_t jb, const LuWork &lw) {
    if (jb <= 1) return RLG_OK;
    // algorithmic bytes: the unit-lower block read, the 64 x 64 inverse written
    RLG_PROF_W("tri_inv", s, ((double)jb * jb / 2 + (double)PNB * PNB) * sizeof(T));
    hipLaunchKernelGGL(k_tri_inv<T>, dim3(1), dim3(256), 0, s, a, (long)lda, (int)j0, (int)jb,
                       linv_of<T>(lw, j0));
    RLG_LAUNCHED();
    return RLG_OK;
}

template <class T>
static int trsm_llu(hipStream_t s, T *a, size_t lda, size_t j0, size_t jb, size_t c0, size_t c1,
                    const LuWork &lw) {
    if (c1 <= c0 || jb <= 1) return RLG_OK;
    // algorithmic bytes: the 64 x 64 inverse and the jb x (c1-c0) block row (read + write)
    RLG_PROF_W("trsm", s, ((double)PNB * PNB + 2.0 * (double)jb * (double)(c1 - c0)) * sizeof(T));
    hipLaunchKernelGGL(k_trsm_llu<T>, dim3(ceil_div(c1 - c0, 64)), dim3(256), 0, s, a, (long)lda,
                       (int)j0, (int)jb, (int)c0, (int)c1, (const T *)linv_of<T>(lw, j0));
    RLG_LAUNCHED();
    return RLG_OK;
}

template <class T>
using UpdateFn = int (*)(hipStream_t, size_t, size_t, size_t, const T *, size_t, const T *, size_t,
                         T *, size_t);

// C -= A * B in working precision.
template <class T>
static int update_native(hipStream_t s, size_t m, size_t n, size_t k, const T *a, size_t lda,
                         const T *b, size_t ldb, T *c, size_t ldc) {
    return gemm_t<T>(s, m, n, k, T(-1), a, lda, b, ldb, T(1), c, ldc);
}

// set by lu_factor_lowp for its duration: precision and the packed-operand
// areas (one for the caller's stream, one for the look-ahead's second stream)
static thread_local int g_lowp_prec = RLG_PREC_F16;
static thread_local void *g_lowp_pack = nullptr;
static thread_local void *g_lowp_pack2 = nullptr;
static thread_local size_t g_lowp_pack_bytes = 0;
static int update_lowp(hipStream_t s, size_t m, size_t n, size_t k, const float *a, size_t lda,
                       const float *b, size_t ldb, float *c, size_t ldc) {
    return gemm_lowp(s, g_lowp_prec, m, n, k, -1.f, a, lda, b, ldb, 1.f, c, ldc, false, false,
                     g_lowp_pack, g_lowp_pack_bytes);
}

// The look-ahead's update on the second stream.  Its workgroups reserve LDS
// they do not use, so that one of them and one workgroup of the panel chain
// (tournament, row swaps, block-row solve, depth-64 update: at most 67 KB)
// share a CU, where two of them would fill its 160 KB: the split FP16
// kernel stages 80 KB, the FP32 one 66 KB.
constexpr size_t LU_SPLIT_PAD = 8 << 10, LU_F32_PAD = 24 << 10;
constexpr size_t LOOKAHEAD_MIN_N = 8192;  // smallest order factored with look-ahead
static int update_lowp_aux(hipStream_t s, size_t m, size_t n, size_t k, const float *a, size_t lda,
                           const float *b, size_t ldb, float *c, size_t ldc) {
    return gemm_lowp(s, g_lowp_prec, m, n, k, -1.f, a, lda, b, ldb, 1.f, c, ldc, false, false,
                     g_lowp_pack2, g_lowp_pack_bytes, LU_SPLIT_PAD);
}
static int update_f32_aux(hipStream_t s, size_t m, size_t n, size_t k, const float *a, size_t lda,
                          const float *b, size_t ldb, float *c, size_t ldc) {
    return gemm_t_padded<float>(s, m, n, k, -1.f, a, lda, b, ldb, 1.f, c, ldc, false, false, LU_F32_PAD);
}

// Events used to order the look-ahead stream against the main stream.
struct StepEvents {
    hipEvent_t ready = nullptr, big = nullptr;
    bool big_pending = false;
    ~StepEvents() {
        if (ready) (void)hipEventDestroy(ready);
        if (big) (void)hipEventDestroy(big);
    }
};

// Factor the outer block of columns [k0, k0+kb): inner panels of PNB columns,
// their swaps on the rest of the block, and the block-local updates.
template <class T>
static int outer_panel(hipStream_t s, size_t n, T *a, size_t lda, int64_t *ipiv, int64_t *info,
                       const LuWork &lw, size_t k0, size_t kb) {
    for (size_t j0 = k0; j0 < k0 + kb; j0 += PNB) {
        const size_t jb = k0 + kb - j0 < (size_t)PNB ? k0 + kb - j0 : PNB;
        bool linv_done = false;
        RLG_TRY(panel_any<T>(s, a, lda, j0, n - j0, jb, ipiv, info, n, lw, &linv_done));
        // L^-1 of the panel's diagonal block, for its two block-row solves
        if (j0 + jb < n && !linv_done) RLG_TRY(tri_inv<T>(s, a, lda, j0, jb, lw));
        // inner swaps on the other columns of the outer block
        RLG_TRY(laswp<T>(s, a, lda, ipiv, j0, jb, k0, j0, j0 + jb, k0 + kb));
        if (j0 + jb < k0 + kb) {
            RLG_TRY(trsm_llu<T>(s, a, lda, j0, jb, j0 + jb, k0 + kb, lw));
            RLG_TRY(update_native<T>(s, n - j0 - jb, k0 + kb - j0 - jb, jb, a + (j0 + jb) * lda + j0,
                                     lda, a + j0 * lda + j0 + jb, lda, a + (j0 + jb) * lda + j0 + jb,
                                     lda));
        }
    }
    return RLG_OK;
}

// Right-looking blocked LU, with look-ahead when `aux_update` is given.
// After outer block k is factored and its block row solved, the trailing
// update is split: the columns of block k+1 are updated first on the
// caller's stream, which then factors block k+1, while the rest of the
// trailing update runs on a second stream (`aux_update`; a stream created
// for the call, default priority).  The next step's row swaps and
// block-row solve wait for that update on an event.  The two streams touch
// disjoint columns meanwhile (block k+1's own panels, swaps and updates stay
// inside its columns), every entry receives the same updates in the same
// order, and the split follows the update's 128-wide tile grid, so the
// factors are the same as without look-ahead.
//
// Look-ahead is used with the tournament panels only.  Their workgroups
// never wait for one another, so they run beside the update on whatever
// CUs are free.  The partial-pivoting panel is a cooperative launch, which
// the runtime does not overlap with other streams' kernels.  An earlier form
// with a non-cooperative exchange panel beside the second stream stalled on
// the GPU.
template <class T>
static int lu_factor_impl(hipStream_t s, size_t n, T *a, size_t lda, int64_t *ipiv, int64_t *info,
                          void *work, UpdateFn<T> big_update, UpdateFn<T> aux_update) {
    // Below this order factorizations are mostly run several at a time (the
    // serving loop's systems go up to order 4096), and those already share
    // the device: there a second stream per factorization measured slower
    // (1,450 -> 975 solves/s at 2,000 requests/s offered).
    LuWork lw = carve(work, n, sizeof(T));
    // exchange tags restart with every factorization
    RLG_TRY(fill_u32(reinterpret_cast<uint32_t *>(lw.xch), 2 * xch_words(sizeof(T)), 0u, s));
    RLG_TRY(fill_u32(reinterpret_cast<uint32_t *>(lw.timeout), 1, 0u, s));
    RLG_TRY(fill_i64(info, 1, (int64_t)n, s));
    if (n == 0) return RLG_OK;
    AuxStream aux;
    hipStream_t s2 = nullptr;
    StepEvents ev;
    if (aux_update && n >= LOOKAHEAD_MIN_N) {
        RLG_TRY(aux.create());
        s2 = aux.s;
        RLG_HIP(hipEventCreateWithFlags(&ev.ready, hipEventDisableTiming));
        RLG_HIP(hipEventCreateWithFlags(&ev.big, hipEventDisableTiming));
    }
    RLG_TRY(outer_panel<T>(s, n, a, lda, ipiv, info, lw, 0, n < (size_t)ONB ? n : ONB));
    for (size_t k0 = 0; k0 < n; k0 += ONB) {
        const size_t kb = n - k0 < (size_t)ONB ? n - k0 : ONB;
        const size_t c0 = k0 + kb;
        if (ev.big_pending) {  // the previous trailing update must be complete
            RLG_HIP(hipStreamWaitEvent(s, ev.big, 0));
            ev.big_pending = false;
        }
        // outer swaps on everything left and right of the outer block
        RLG_TRY(laswp<T>(s, a, lda, ipiv, k0, kb, 0, k0, c0, n));
        if (c0 >= n) break;
        for (size_t j0 = k0; j0 < c0; j0 += PNB) {
            const size_t jb = c0 - j0 < (size_t)PNB ? c0 - j0 : PNB;
            RLG_TRY(trsm_llu<T>(s, a, lda, j0, jb, c0, n, lw));
            if (j0 + jb < c0)
                RLG_TRY(update_native<T>(s, c0 - j0 - jb, n - c0, jb, a + (j0 + jb) * lda + j0, lda,
                                         a + j0 * lda + c0, lda, a + (j0 + jb) * lda + c0, lda));
        }
        const size_t kb2 = n - c0 < (size_t)ONB ? n - c0 : ONB;
        if (!s2 || c0 + kb2 >= n) {
            RLG_TRY(big_update(s, n - c0, n - c0, kb, a + c0 * lda + k0, lda, a + k0 * lda + c0, lda,
                               a + c0 * lda + c0, lda));
        } else {
            // look-ahead columns first, then the rest on the second stream
            RLG_TRY(big_update(s, n - c0, kb2, kb, a + c0 * lda + k0, lda, a + k0 * lda + c0, lda,
                               a + c0 * lda + c0, lda));
            RLG_HIP(hipEventRecord(ev.ready, s));
            RLG_HIP(hipStreamWaitEvent(s2, ev.ready, 0));
            RLG_TRY(aux_update(s2, n - c0, n - c0 - kb2, kb, a + c0 * lda + k0, lda,
                               a + k0 * lda + c0 + kb2, lda, a + c0 * lda + c0 + kb2, lda));
            RLG_HIP(hipEventRecord(ev.big, s2));
            ev.big_pending = true;
        }
        RLG_TRY(outer_panel<T>(s, n, a, lda, ipiv, info, lw, c0, kb2));
    }
    if (ev.big_pending) RLG_HIP(hipStreamWaitEvent(s, ev.big, 0));
    return RLG_OK;
}

template <class T>
int lu_factor(hipStream_t s, size_t n, T *a, size_t lda, int64_t *ipiv, int64_t *info, void *work) {
    return lu_factor_impl<T>(s, n, a, lda, ipiv, info, work, &update_native<T>, nullptr);
}
template int lu_factor<double>(hipStream_t, size_t, double *, size_t, int64_t *, int64_t *, void *);
template int lu_factor<float>(hipStream_t, size_t, float *, size_t, int64_t *, int64_t *, void *);

// The trailing updates are at most n x n with depth ONB; the look-ahead's
// second stream has a packed-operand area of its own.
size_t lu_lowp_work_bytes(size_t n) {
    return align_up(lu_work_bytes(n, 4)) + 2 * align_up(gemm_lowp_pack_bytes(n, n, ONB)) +
           tslu_work_bytes(n);
}

int lu_factor_lowp(hipStream_t s, int prec, int pivot, size_t n, float *a, size_t lda, int64_t *ipiv,
                   int64_t *info, void *work) {
    char *w = static_cast<char *>(work);
    const size_t pack = align_up(gemm_lowp_pack_bytes(n, n, ONB));
    g_lowp_prec = prec;
    g_lowp_pack = w + align_up(lu_work_bytes(n, 4));
    g_lowp_pack2 = static_cast<char *>(g_lowp_pack) + pack;
    g_lowp_pack_bytes = gemm_lowp_pack_bytes(n, n, ONB);
    g_tslu_work = nullptr;
    g_panel_exact = false;
    int st = RLG_OK;
    UpdateFn<float> aux = nullptr;
    if (pivot == RLG_PIVOT_TOURNAMENT) {
        g_tslu_work = static_cast<char *>(g_lowp_pack2) + pack;
        st = tslu_reset(s, g_tslu_work, n);
        aux = prec == RLG_PREC_F32 ? &update_f32_aux : &update_lowp_aux;
    }
    if (st == RLG_OK)
        st = lu_factor_impl<float>(s, n, a, lda, ipiv, info, work,
                                   prec == RLG_PREC_F32 ? &update_native<float> : &update_lowp, aux);
    g_lowp_pack = nullptr;
    g_lowp_pack2 = nullptr;
    g_lowp_pack_bytes = 0;
    g_tslu_work = nullptr;
    g_panel_exact = true;
    return st;
}

// ---------------------------------------------------------------- solve

// tmp[r][k] = b[r][raw[k]]: y = P b with (P b)[k] = b[raw[k]], which is the
// reference's permute_vector_into_buffer (permutation_matrix.rs:369-382)
// applied with the stored inverse permutation.
template <class T>
__global__ void k_gather_rows(size_t n, size_t nrhs, const int64_t *__restrict__ raw,
                              const T *__restrict__ b, size_t ldb, T *__restrict__ y, size_t ldy) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t r = blockIdx.y;
    if (i < n && r < nrhs) y[r * ldy + i] = b[r * ldb + raw[i]];
}

template <class T>
int lu_solve(hipStream_t s, size_t n, const T *lu, size_t lda, const int64_t *raw, size_t nrhs,
             const T *b, size_t ldb, T *x, size_t ldx, T *tmp) {
    if (n == 0 || nrhs == 0) return RLG_OK;
    {
        RLG_PROF("gather", s);
        hipLaunchKernelGGL(k_gather_rows<T>, dim3(ceil_div(n, 256), (unsigned)nrhs), dim3(256), 0, s, n,
                           nrhs, raw, b, ldb, tmp, n);
        RLG_LAUNCHED();
    }
    // y = L^-1 (P b) lands in x (b is no longer read), then U^-1 y in tmp
    RLG_TRY(trsv_lower<T>(s, n, lu, lda, true, nrhs, tmp, n, x, ldx));
    RLG_TRY(trsv_upper<T>(s, n, lu, lda, nrhs, x, ldx, tmp, n));
    RLG_HIP(hipMemcpy2DAsync(x, ldx * sizeof(T), tmp, n * sizeof(T), n * sizeof(T), nrhs,
                             hipMemcpyDeviceToDevice, s));
    return RLG_OK;
}
template int lu_solve<double>(hipStream_t, size_t, const double *, size_t, const int64_t *, size_t,
                              const double *, size_t, double *, size_t, double *);
template int lu_solve<float>(hipStream_t, size_t, const float *, size_t, const int64_t *, size_t,
                             const float *, size_t, float *, size_t, float *);

}  // namespace rlg
