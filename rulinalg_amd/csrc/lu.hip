// Right-looking blocked LU with partial pivoting (PartialPivLu::decompose,
// src/matrix/decomposition/lu.rs:163-194) and its solve (lu.rs:231-244).
//
// Blocking: outer blocks of ONB = 256 columns drive the trailing Schur
// update as one MFMA GEMM of depth 256; inside an outer block, inner panels
// of PNB = 64 columns are factored by a cooperative kernel and applied to
// the rest of the outer block by a 64-row triangular solve and a depth-64
// GEMM.
//
// The panel kernel keeps the whole (n - k) x 64 panel resident in LDS,
// spread over G workgroups of PR rows each.  Every column costs one
// grid-wide barrier: each workgroup publishes its local pivot candidate
// (|value|, row, and the full 64-wide row) and the content of the current
// diagonal row; after the barrier every workgroup reduces the G candidates
// to the same winner, performs its part of the row swap, and eliminates its
// rows.  The pivot rule is the reference's: the first row of maximal
// magnitude at or below the diagonal (strict `>` scan, lu.rs:170-178); the
// multiplier is a division and the update a separately rounded multiply and
// subtract (lu.rs:603-616), so a matrix no wider than one panel is factored
// bit-identically to the reference.
#include <climits>
#include <type_traits>

#include "kernels.hpp"

namespace rlg {

constexpr int PNB = 64;    // inner panel width
// Panel rows per workgroup (LDS-resident): 256 (133 KB of LDS in FP64, one
// workgroup per CU).  A grid barrier costs about 85 ns per participating
// workgroup on MI355X, so fewer, taller workgroups make every column cheaper.
constexpr int PR = 256;
template <class T> struct PanelRows { static constexpr int value = PR; };
template <> struct PanelRows<float> { static constexpr int value = 2 * PR; };  // same 133 KB
constexpr int ONB = 256;   // outer block: depth of the trailing GEMM
constexpr int PTHREADS = 256;
// Barrier counters per panel: top at 0, group i (0..7) at 32 (i + 1), i.e.
// up to index 256, so each panel's counters span 257 entries (128-byte
// spacing); the next panel's start at BAR_STRIDE.
constexpr int BAR_STRIDE = 512;
static_assert(32 * 8 < BAR_STRIDE, "group counters must stay inside the panel's block");

template <class T> __device__ __forceinline__ T eps_of();
template <> __device__ __forceinline__ double eps_of<double>() { return 2.220446049250313e-16; }
template <> __device__ __forceinline__ float eps_of<float>() { return 1.1920929e-7f; }

// Workspace layout (bytes): barrier counters, candidate slots, swap map.
struct LuWork {
    unsigned *bar;     // BAR_STRIDE barrier counters per inner panel
    void *cval;        // [2][GMAX] T
    int *crow;         // [2][GMAX]
    void *cdat;        // [2][GMAX][PNB] T
    void *rowj;        // [2][PNB] T
    int *map;          // [1 + 2 * 2 * ONB]: count, dst[], src[]
    int *timeout;      // barrier timeout flag
};
constexpr int GMAX = 1024;

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

size_t lu_work_bytes(size_t n, size_t elem) {
    size_t panels = (n + PNB - 1) / PNB + 1;
    return align_up(panels * BAR_STRIDE * 4) + align_up(2 * GMAX * elem) + align_up(2 * GMAX * 4) +
           align_up(2 * GMAX * PNB * elem) + align_up(2 * PNB * elem) +
           align_up((1 + 4 * ONB) * 4) + align_up(4);
}

static LuWork carve(void *work, size_t n, size_t elem) {
    char *p = static_cast<char *>(work);
    LuWork w;
    size_t panels = (n + PNB - 1) / PNB + 1;
    w.bar = reinterpret_cast<unsigned *>(p); p += align_up(panels * BAR_STRIDE * 4);
    w.cval = p; p += align_up(2 * GMAX * elem);
    w.crow = reinterpret_cast<int *>(p); p += align_up(2 * GMAX * 4);
    w.cdat = p; p += align_up(2 * GMAX * PNB * elem);
    w.rowj = p; p += align_up(2 * PNB * elem);
    w.map = reinterpret_cast<int *>(p); p += align_up((1 + 4 * ONB) * 4);
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

// Candidate order: larger magnitude first; equal magnitudes go to the lower
// row index, which reproduces the reference's first-maximum scan.
template <class T>
__device__ __forceinline__ bool better(T v2, int i2, T v1, int i1) {
    if (i2 == INT_MAX) return false;
    if (i1 == INT_MAX) return true;
    if (v2 > v1) return true;
    if (v1 > v2) return false;
    return i2 < i1;
}

template <class T>
__device__ __forceinline__ void wave_argmax(T &v, int &i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(i, off, 64);
        if (better(ov, oi, v, i)) { v = ov; i = oi; }
    }
}

// Grid barrier for the panel's workgroups, the epoch-th of this panel.
// Two-level arrival: with more than 8 workgroups, workgroup g first meets
// the others of its group g % 8 (the workgroups a round-robin dispatch puts
// on one XCD) on a group counter, and only the last of each group arrives at
// the top counter, so no counter takes more than max(8, G/8) atomic adds
// per column; everyone then waits on the top counter.  (A flat counter costs
// about 85 ns per arriving workgroup.)  Release/acquire: each workgroup's
// release fence precedes its add; the last arriver of a group acquires
// (through the chain of adds on the group counter) and releases again before
// its add on the top counter; waiters acquire after reading it.  Bounded:
// after 2^22 polls the shared timeout flag is raised; every waiter also
// leaves as soon as it sees the flag, so a barrier that cannot complete ends
// the kernel (with an error the host reports) instead of hanging it.
__device__ __forceinline__ void grid_barrier(unsigned *bar, int G, unsigned epoch, int *timeout) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned target;
        if (G <= 8) {
            __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            target = (unsigned)G * epoch;
        } else {
            const int grp = blockIdx.x & 7;
            const unsigned ng = (unsigned)((G - grp + 7) >> 3);
            const unsigned old =
                __hip_atomic_fetch_add(bar + 32 * (grp + 1), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1 == ng * epoch) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            target = 8u * epoch;
        }
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (++spins > (1u << 22)) {
                __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Factor the panel A[k:k+m, k:k+w] (w <= PNB) in place.
template <class T>
__global__ __launch_bounds__(PTHREADS) void k_lu_panel(T *__restrict__ A, long lda, int k, int m,
                                                       int w, int64_t *__restrict__ ipiv,
                                                       int64_t *__restrict__ info, int n,
                                                       unsigned *bar, T *cval, int *crow,
                                                       T *cdat, T *rowj, int *timeout) {
    constexpr int PRT = PanelRows<T>::value;
    __shared__ T tile[PRT][PNB + 1];
    __shared__ T prow[PNB];
    __shared__ T red_v[PTHREADS / 64];
    __shared__ int red_i[PTHREADS / 64];
    __shared__ int s_piv;

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
            if (r < rows && c < w) tile[r][c] = v[q];
        }
    }
    __syncthreads();

    for (int j = 0; j < w; ++j) {
        const int par = j & 1;
        // 1. local candidate: first maximum of |tile[r][j]| over owned rows >= j
        T bv = T(0);
        int bi = INT_MAX;
        for (int r = tid; r < rows; r += PTHREADS) {
            const int gr = r0 + r;
            if (gr >= j) {
                T v = fabs(tile[r][j]);
                if (better(v, gr, bv, bi)) { bv = v; bi = gr; }
            }
        }
        wave_argmax(bv, bi);
        if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; }
        __syncthreads();
        if (tid == 0) {
            for (int q = 1; q < PTHREADS / 64; ++q)
                if (better(red_v[q], red_i[q], bv, bi)) { bv = red_v[q]; bi = red_i[q]; }
            red_v[0] = bv;
            red_i[0] = bi;
        }
        __syncthreads();
        const int li = red_i[0];
        const bool own_j = (j >= r0 && j < r0 + rows);

        int p;
        if (G > 1) {
            // 2. publish candidate row and the diagonal row, then meet
            T *my_cdat = cdat + ((long)par * G + g) * PNB;
            if (tid == 0) { cval[par * G + g] = red_v[0]; crow[par * G + g] = li; }
            if (li != INT_MAX && tid < w) my_cdat[tid] = tile[li - r0][tid];
            if (own_j && tid < w) rowj[par * PNB + tid] = tile[j - r0][tid];
            grid_barrier(bar, G, (unsigned)(j + 1), timeout);
            // 3. every workgroup reduces all candidates to the same winner
            if (wave == 0) {
                T v = T(0);
                int i = INT_MAX, gw = 0;
                for (int q = lane; q < G; q += 64) {
                    T qv = cval[par * G + q];
                    int qi = crow[par * G + q];
                    if (better(qv, qi, v, i)) { v = qv; i = qi; gw = q; }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    T ov = __shfl_xor(v, off, 64);
                    int oi = __shfl_xor(i, off, 64);
                    int og = __shfl_xor(gw, off, 64);
                    if (better(ov, oi, v, i)) { v = ov; i = oi; gw = og; }
                }
                if (lane == 0) s_piv = i;
                // winner's row data
                if (lane < w) prow[lane] = cdat[((long)par * G + gw) * PNB + lane];
            }
            __syncthreads();
            p = s_piv;
            // 4. swap rows j and p in the owning tiles (old row j comes from
            //    the published copy when it lives in another workgroup)
            if (p != j && tid < w) {
                const T old_j = own_j ? tile[j - r0][tid] : rowj[par * PNB + tid];
                if (own_j) tile[j - r0][tid] = prow[tid];
                if (p >= r0 && p < r0 + rows) tile[p - r0][tid] = old_j;
            }
        } else {
            p = li;
            if (tid < w) {
                const T vp = tile[p][tid];
                const T vj = tile[j][tid];
                prow[tid] = vp;
                if (p != j) { tile[j][tid] = vp; tile[p][tid] = vj; }
            }
        }
        __syncthreads();
        if (g == 0 && tid == 0) {
            ipiv[k + j] = (int64_t)(k + p);
            if (fabs(prow[j]) < eps_of<T>() && *info == (int64_t)n) *info = k + j;
        }

        // 5. eliminate: mult = a[i][j] / piv; a[i][c] = a[i][c] - mult * piv_row[c]
        {
#pragma clang fp contract(off)
            const T piv = prow[j];
            constexpr int RPT = PRT / PTHREADS;  // panel rows per thread
            static_assert(RPT * PTHREADS == PRT, "whole rows per thread");
            T mult[RPT];
            bool act[RPT];
#pragma unroll
            for (int h = 0; h < RPT; ++h) {
                const int r = tid + h * PTHREADS;
                act[h] = r < rows && r0 + r > j;
                mult[h] = act[h] ? tile[r][j] / piv : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < RPT; ++h) {
                const int r = tid + h * PTHREADS;
                if (act[h]) {
                    tile[r][j] = mult[h];
                    // eight columns per uniform branch, their LDS reads issued
                    // together (a plain loop waits one LDS latency per column)
#pragma unroll
                    for (int c0 = 0; c0 < PNB; c0 += 8) {
                        if (c0 + 7 > j && c0 < w) {
#pragma unroll
                            for (int q = 0; q < 8; ++q) {
                                const int c = c0 + q;
                                if (c > j && c < w) tile[r][c] = tile[r][c] - mult[h] * prow[c];
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    for (int e = tid; e < rows * w; e += PTHREADS) {
        const int r = e / w, c = e - r * w;
        A[(long)(k + r0 + r) * lda + k + c] = tile[r][c];
    }
}

// ---------------------------------------------------------------- row swaps

// Net effect of the swaps ipiv[k0 .. k0+cnt) as a list of (dst, src) rows:
// after the swaps, row dst holds what row src held before.  One wave walks
// the swaps in order with every table in registers (entry e of a table at
// lane e % 64, slot e / 64, read with readlane): `cont` holds what position
// k0 + e now contains; rows outside [k0, k0+cnt) that get touched are kept
// in (erow, esrc) and found with a ballot.  The list is then written with a
// ballot prefix count.
constexpr int SMS = ONB / 64;  // table slots per lane

__device__ __forceinline__ int tab_get(const int (&t)[SMS], int e) {
    int v = t[0];
#pragma unroll
    for (int i = 1; i < SMS; ++i)
        if ((e >> 6) == i) v = t[i];
    return __builtin_amdgcn_readlane(v, e & 63);
}
__device__ __forceinline__ void tab_set(int (&t)[SMS], int e, int v) {
    if ((int)threadIdx.x == (e & 63)) {
#pragma unroll
        for (int i = 0; i < SMS; ++i)
            if ((e >> 6) == i) t[i] = v;
    }
}

__global__ __launch_bounds__(64) void k_swap_map(const int64_t *__restrict__ ipiv, int k0, int cnt,
                                                 int *__restrict__ map) {
    const int lane = threadIdx.x;
    int cont[SMS], pv[SMS], erow[SMS], esrc[SMS];
#pragma unroll
    for (int i = 0; i < SMS; ++i) {
        const int e = i * 64 + lane;
        cont[i] = k0 + e;
        pv[i] = e < cnt ? (int)ipiv[k0 + e] : k0 + e;
        erow[i] = -1;
        esrc[i] = -1;
    }
    int n_ext = 0;  // uniform
    for (int s = 0; s < cnt; ++s) {
        const int p = tab_get(pv, s);
        if (p == k0 + s) continue;
        if (p < k0 + cnt) {
            const int a = tab_get(cont, s), b = tab_get(cont, p - k0);
            tab_set(cont, s, b);
            tab_set(cont, p - k0, a);
        } else {
            int idx = -1;
#pragma unroll
            for (int i = 0; i < SMS; ++i) {
                const unsigned long long bal = __ballot(erow[i] == p);
                if (idx < 0 && bal) idx = i * 64 + __ffsll((long long)bal) - 1;
            }
            if (idx < 0) {
                idx = n_ext++;
                tab_set(erow, idx, p);
                tab_set(esrc, idx, p);
            }
            const int a = tab_get(cont, s), e = tab_get(esrc, idx);
            tab_set(cont, s, e);
            tab_set(esrc, idx, a);
        }
    }
    // the list: positions whose content changed, then touched outside rows
    int c = 0;
#pragma unroll
    for (int i = 0; i < SMS; ++i) {
        const int e = i * 64 + lane;
        const bool mv = e < cnt && cont[i] != k0 + e;
        const unsigned long long bal = __ballot(mv);
        if (mv) {
            const int at = c + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
            map[1 + at] = k0 + e;
            map[1 + 2 * ONB + at] = cont[i];
        }
        c += __popcll(bal);
    }
#pragma unroll
    for (int i = 0; i < SMS; ++i) {
        const int e = i * 64 + lane;
        const bool mv = e < n_ext && esrc[i] != erow[i];
        const unsigned long long bal = __ballot(mv);
        if (mv) {
            const int at = c + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
            map[1 + at] = erow[i];
            map[1 + 2 * ONB + at] = esrc[i];
        }
        c += __popcll(bal);
    }
    if (lane == 0) map[0] = c;
}

// Applies a swap map to columns [c0, c1) and [c2, c3).  Each workgroup owns
// LCOLS columns: it reads every source row of the map into LDS, then writes
// the destinations, so the in-place permutation needs no ordering.
constexpr int LCOLS = 32;
template <class T>
__global__ __launch_bounds__(256) void k_laswp(T *__restrict__ A, long lda, const int *__restrict__ map,
                                               int c0, int c1, int c2, int c3) {
    __shared__ T stage[2 * ONB][LCOLS];
    __shared__ int dst[2 * ONB], src[2 * ONB];
    const int cnt = map[0];
    if (cnt == 0) return;
    for (int i = threadIdx.x; i < cnt; i += 256) { dst[i] = map[1 + i]; src[i] = map[1 + 2 * ONB + i]; }
    // column block -> absolute columns
    const int w1 = c1 - c0;
    const int cb = blockIdx.x * LCOLS;
    __syncthreads();
    const int cl = threadIdx.x % LCOLS, rg = threadIdx.x / LCOLS;  // 8 row groups
    const int lc = cb + cl;
    const int col = lc < w1 ? c0 + lc : c2 + (lc - w1);
    const bool okc = lc < w1 + (c3 - c2);
    // eight loads in flight per thread before their LDS stores
    constexpr int RG = 256 / LCOLS;
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
    for (int i = rg; i < cnt; i += 256 / LCOLS)
        if (okc) A[(long)dst[i] * lda + col] = stage[i][cl];
}

// ---------------------------------------------------------------- TRSM

// B := L^-1 B for the unit-lower jb x jb block L = A[j0:j0+jb, j0:j0+jb]
// (jb <= 64) and B = A[j0:j0+jb, c0:c1), in place, column sweep:
// for each step c, b[r] = b[r] - l[r][c] * b[c] for r > c.  This is the
// order in which gaussian_elimination (lu.rs:603-616) updates the rows of
// the block, with the same separately rounded multiply and subtract.
//
// L and the 64-column slab of B are staged in LDS with every load issued
// before the stores; four waves split the rows of each sweep step.
template <class T>
__global__ __launch_bounds__(256) void k_trsm_llu(T *__restrict__ A, long lda, int j0, int jb, int c0,
                                                  int c1) {
    __shared__ T L[PNB][PNB + 1];
    __shared__ T X[PNB][65];
    const int tid = threadIdx.x;
    const int cb = c0 + blockIdx.x * 64;
    const int nc = min(64, c1 - cb);
    {
        constexpr int PER = PNB * PNB / 256;
        T vl[PER], vx[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e >> 6, c = e & 63;
            vl[q] = (r < jb && c < jb) ? A[(long)(j0 + r) * lda + j0 + c] : T(0);
            vx[q] = (r < jb && c < nc) ? A[(long)(j0 + r) * lda + cb + c] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e >> 6, c = e & 63;
            L[r][c] = vl[q];
            X[r][c] = vx[q];
        }
    }
    __syncthreads();
    // each wave owns 16 whole columns (lane = column + 16 * row group), so a
    // column's steps only depend on LDS writes of the same wave, which the
    // LDS applies in order: no workgroup barrier inside the sweep
    const int cj = (tid >> 6) * 16 + (tid & 15), rq = (tid & 63) >> 4;
    {
#pragma clang fp contract(off)
        for (int c = 0; c < jb - 1; ++c) {
            const T xc = X[c][cj];
            // the (up to 16) rows of this lane's group, unrolled so their LDS
            // reads are issued together
#pragma unroll
            for (int i = 0; i < PNB / 4; ++i) {
                const int r = c + 1 + rq + 4 * i;
                if (r < jb) X[r][cj] = X[r][cj] - L[r][c] * xc;
            }
            __builtin_amdgcn_wave_barrier();
            __asm__ __volatile__("" ::: "memory");  // keep step c + 1's reads after these writes
        }
    }
    __syncthreads();
    for (int e = tid; e < jb * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (c < nc) A[(long)(j0 + r) * lda + cb + c] = X[r][c];
    }
}

// ---------------------------------------------------------------- driver

// Set by lu_factor_lowp for its duration: FP32 panels are then factored with
// tournament pivoting (tslu.hip) in this workspace, one epoch per panel.
static thread_local void *g_tslu_work = nullptr;
static thread_local unsigned g_tslu_epoch = 0;

template <class T>
static int panel(hipStream_t s, T *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
                 int64_t *info, size_t n, const LuWork &lw, unsigned *bar);

template <class T>
static int panel_any(hipStream_t s, T *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
                     int64_t *info, size_t n, const LuWork &lw, unsigned *bar) {
    if constexpr (std::is_same<T, float>::value) {
        if (g_tslu_work && m <= tslu_max_rows())
            return tslu_panel(s, a, lda, k, m, w, ipiv, info, n, g_tslu_work, ++g_tslu_epoch,
                              lw.timeout);
    }
    return panel<T>(s, a, lda, k, m, w, ipiv, info, n, lw, bar);
}

template <class T>
static int panel(hipStream_t s, T *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
                 int64_t *info, size_t n, const LuWork &lw, unsigned *bar) {
    const unsigned G = ceil_div(m, PanelRows<T>::value);
    if (G > GMAX) {
        set_error("LU panel of %zu rows needs %u workgroups (max %d)", m, G, GMAX);
        return RLG_E_INVALID_ARG;
    }
    // algorithmic bytes: the m x w panel read once and written once
    RLG_PROF_W("lu_panel", s, 2.0 * (double)m * (double)w * sizeof(T));
    hipLaunchKernelGGL(k_lu_panel<T>, dim3(G), dim3(PTHREADS), 0, s, a, (long)lda, (int)k, (int)m,
                       (int)w, ipiv, info, (int)n, bar, static_cast<T *>(lw.cval), lw.crow,
                       static_cast<T *>(lw.cdat), static_cast<T *>(lw.rowj), lw.timeout);
    RLG_LAUNCHED();
    return RLG_OK;
}

template <class T>
static int laswp(hipStream_t s, T *a, size_t lda, const int64_t *ipiv, size_t k0, size_t cnt,
                 size_t c0, size_t c1, size_t c2, size_t c3, const LuWork &lw) {
    const size_t ncols = (c1 - c0) + (c3 - c2);
    if (ncols == 0 || cnt == 0) return RLG_OK;
    // algorithmic bytes: at most 2 cnt rows change, each read and written
    RLG_PROF_W("laswp", s, 2.0 * 2.0 * (double)cnt * (double)ncols * sizeof(T));
    hipLaunchKernelGGL(k_swap_map, dim3(1), dim3(64), 0, s, ipiv, (int)k0, (int)cnt, lw.map);
    RLG_LAUNCHED();
    hipLaunchKernelGGL(k_laswp<T>, dim3(ceil_div(ncols, LCOLS)), dim3(256), 0, s, a, (long)lda,
                       lw.map, (int)c0, (int)c1, (int)c2, (int)c3);
    RLG_LAUNCHED();
    return RLG_OK;
}

template <class T>
static int trsm_llu(hipStream_t s, T *a, size_t lda, size_t j0, size_t jb, size_t c0, size_t c1) {
    if (c1 <= c0 || jb <= 1) return RLG_OK;
    // algorithmic bytes: the unit-lower jb x jb block and the jb x (c1-c0) block row (read + write)
    RLG_PROF_W("trsm", s, ((double)jb * jb / 2 + 2.0 * (double)jb * (double)(c1 - c0)) * sizeof(T));
    hipLaunchKernelGGL(k_trsm_llu<T>, dim3(ceil_div(c1 - c0, 64)), dim3(256), 0, s, a, (long)lda,
                       (int)j0, (int)jb, (int)c0, (int)c1);
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

// set by lu_factor_lowp for its duration: precision and the packed-operand area
static thread_local int g_lowp_prec = RLG_PREC_F16;
static thread_local void *g_lowp_pack = nullptr;
static thread_local size_t g_lowp_pack_bytes = 0;
static int update_lowp(hipStream_t s, size_t m, size_t n, size_t k, const float *a, size_t lda,
                       const float *b, size_t ldb, float *c, size_t ldc) {
    return gemm_lowp(s, g_lowp_prec, m, n, k, -1.f, a, lda, b, ldb, 1.f, c, ldc, false, false,
                     g_lowp_pack, g_lowp_pack_bytes);
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
        RLG_TRY(panel_any<T>(s, a, lda, j0, n - j0, jb, ipiv, info, n, lw, lw.bar + (j0 / PNB) * BAR_STRIDE));
        // inner swaps on the other columns of the outer block
        RLG_TRY(laswp<T>(s, a, lda, ipiv, j0, jb, k0, j0, j0 + jb, k0 + kb, lw));
        if (j0 + jb < k0 + kb) {
            RLG_TRY(trsm_llu<T>(s, a, lda, j0, jb, j0 + jb, k0 + kb));
            RLG_TRY(update_native<T>(s, n - j0 - jb, k0 + kb - j0 - jb, jb, a + (j0 + jb) * lda + j0,
                                     lda, a + j0 * lda + j0 + jb, lda, a + (j0 + jb) * lda + j0 + jb,
                                     lda));
        }
    }
    return RLG_OK;
}

// Right-looking blocked LU with look-ahead.  After outer block k is factored
// and its block row solved, the trailing update is split: the columns of
// block k+1 are updated first on the main stream, which then factors block
// k+1 while the rest of the trailing update runs on the auxiliary stream.
// The next step's row swaps and block-row solve wait for that update.
//
// LOOK_AHEAD is off: a run with it on stalled on the GPU and the stall has
// not been traced to its cause, so every update runs on the main stream.
constexpr bool LOOK_AHEAD = false;
// The packed low-precision operand area serves one trailing update at a time.
static_assert(!LOOK_AHEAD, "concurrent trailing updates would share the packed operand area");
template <class T>
static int lu_factor_impl(hipStream_t s, size_t n, T *a, size_t lda, int64_t *ipiv, int64_t *info,
                          void *work, UpdateFn<T> big_update) {
    LuWork lw = carve(work, n, sizeof(T));
    const size_t panels = (n + PNB - 1) / PNB + 1;
    RLG_TRY(fill_u32(lw.bar, panels * BAR_STRIDE, 0u, s));
    RLG_TRY(fill_u32(reinterpret_cast<uint32_t *>(lw.timeout), 1, 0u, s));
    RLG_TRY(fill_i64(info, 1, (int64_t)n, s));
    if (n == 0) return RLG_OK;
    hipStream_t s2 = nullptr;
    StepEvents ev;
    if (LOOK_AHEAD && n > 2 * (size_t)ONB) {
        RLG_TRY(aux_stream(&s2));
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
        RLG_TRY(laswp<T>(s, a, lda, ipiv, k0, kb, 0, k0, c0, n, lw));
        if (c0 >= n) break;
        for (size_t j0 = k0; j0 < c0; j0 += PNB) {
            const size_t jb = c0 - j0 < (size_t)PNB ? c0 - j0 : PNB;
            RLG_TRY(trsm_llu<T>(s, a, lda, j0, jb, c0, n));
            if (j0 + jb < c0)
                RLG_TRY(update_native<T>(s, c0 - j0 - jb, n - c0, jb, a + (j0 + jb) * lda + j0, lda,
                                         a + j0 * lda + c0, lda, a + (j0 + jb) * lda + c0, lda));
        }
        const size_t kb2 = n - c0 < (size_t)ONB ? n - c0 : ONB;
        if (!s2 || c0 + kb2 >= n) {
            RLG_TRY(big_update(s, n - c0, n - c0, kb, a + c0 * lda + k0, lda, a + k0 * lda + c0, lda,
                               a + c0 * lda + c0, lda));
        } else {
            // look-ahead columns first, then the rest on the auxiliary stream
            RLG_TRY(big_update(s, n - c0, kb2, kb, a + c0 * lda + k0, lda, a + k0 * lda + c0, lda,
                               a + c0 * lda + c0, lda));
            RLG_HIP(hipEventRecord(ev.ready, s));
            RLG_HIP(hipStreamWaitEvent(s2, ev.ready, 0));
            RLG_TRY(big_update(s2, n - c0, n - c0 - kb2, kb, a + c0 * lda + k0, lda,
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
    return lu_factor_impl<T>(s, n, a, lda, ipiv, info, work, &update_native<T>);
}
template int lu_factor<double>(hipStream_t, size_t, double *, size_t, int64_t *, int64_t *, void *);
template int lu_factor<float>(hipStream_t, size_t, float *, size_t, int64_t *, int64_t *, void *);

// The trailing updates are at most n x n with depth ONB.
size_t lu_lowp_work_bytes(size_t n) {
    return align_up(lu_work_bytes(n, 4)) + align_up(gemm_lowp_pack_bytes(n, n, ONB)) +
           tslu_work_bytes();
}

int lu_factor_lowp(hipStream_t s, int prec, int pivot, size_t n, float *a, size_t lda, int64_t *ipiv,
                   int64_t *info, void *work) {
    char *w = static_cast<char *>(work);
    g_lowp_prec = prec;
    g_lowp_pack = w + align_up(lu_work_bytes(n, 4));
    g_lowp_pack_bytes = gemm_lowp_pack_bytes(n, n, ONB);
    g_tslu_work = nullptr;
    g_tslu_epoch = 0;
    int st = RLG_OK;
    if (pivot == RLG_PIVOT_TOURNAMENT) {
        g_tslu_work = w + align_up(lu_work_bytes(n, 4)) + align_up(g_lowp_pack_bytes);
        st = tslu_reset(s, g_tslu_work);
    }
    if (st == RLG_OK)
        st = lu_factor_impl<float>(s, n, a, lda, ipiv, info, work,
                                   prec == RLG_PREC_F32 ? &update_native<float> : &update_lowp);
    g_lowp_pack = nullptr;
    g_lowp_pack_bytes = 0;
    g_tslu_work = nullptr;
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
