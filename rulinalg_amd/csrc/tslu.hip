// Tournament pivoting (TSLU, the panel step of CALU: Grigori, Demmel and
// Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)) for the FP32 panels of the
// mixed-precision LU.  The plain FP64 drop-in keeps the reference's partial
// pivoting (lu.hip); the mixed factorization only has to be stable enough for
// iterative refinement, which TSLU is in practice (its growth factor behaves
// like partial pivoting's), and it needs a handful of synchronizations per
// panel instead of one grid-wide barrier per column.
//
// One launch factors the m x w panel A[k:k+m, k:k+w] (w <= 64), one wave per
// 128-row chunk:
//   1. every wave runs partial pivoting on its chunk (rows in registers, two
//      per lane) and publishes the 64 winning rows (their original values)
//      and their row numbers;
//   2. a binary tree merges pairs of candidate sets the same way: wave g
//      waits only for its partner g + 2^(l-1) (a point-to-point flag);
//   3. wave 0 ends with the w pivot rows in order and their factored block
//      [L11 \ U11]; it turns the pivot rows into the sequential swaps
//      ipiv[k + j] (the reference's swap semantics), writes the block, the
//      inverse of U11 and the list of rows whose position changes, and
//      raises the `done` flag;
//   4. every wave computes L21 = A21 * inv(U11) for its non-pivot rows and
//      writes each to its final row.
// Every wait is bounded like the panel barrier in lu.hip (2^22 polls, then
// the shared timeout flag, which the host turns into an error).
#include <climits>

#include "kernels.hpp"

namespace rlg {

constexpr int TW = 64;      // panel width
constexpr int TR = 128;     // rows per chunk (two per lane)
constexpr int TMAXG = 256;  // chunks: panels of up to TMAXG * TR rows
constexpr int TLEV = 16;    // flag values per panel epoch (levels 0..8, done = 15)

struct TsluWork {
    unsigned *flag;  // [TMAXG]: epoch * TLEV + level published
    unsigned *done;  // [1]
    int *ids;        // [TMAXG][TW] candidate row numbers
    float *rows;     // [TMAXG][TW][TW] candidate rows, original values
    float *F;        // [TW][TW] factored pivot block
    float *Uinv;     // [TW][TW] inverse of U11
    int *map;        // [1 + 2 * 2 * TW]: count, rows[], final positions[]
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

size_t tslu_work_bytes() {
    return al(TMAXG * 4) + al(4) + al((size_t)TMAXG * TW * 4) + al((size_t)TMAXG * TW * TW * 4) +
           al(TW * TW * 4) + al(TW * TW * 4) + al((1 + 4 * TW) * 4);
}

static TsluWork tslu_carve(void *work) {
    char *p = static_cast<char *>(work);
    TsluWork t;
    t.flag = reinterpret_cast<unsigned *>(p); p += al(TMAXG * 4);
    t.done = reinterpret_cast<unsigned *>(p); p += al(4);
    t.ids = reinterpret_cast<int *>(p); p += al((size_t)TMAXG * TW * 4);
    t.rows = reinterpret_cast<float *>(p); p += al((size_t)TMAXG * TW * TW * 4);
    t.F = reinterpret_cast<float *>(p); p += al(TW * TW * 4);
    t.Uinv = reinterpret_cast<float *>(p); p += al(TW * TW * 4);
    t.map = reinterpret_cast<int *>(p);
    return t;
}

size_t tslu_max_rows() { return (size_t)TMAXG * TR; }

int tslu_reset(hipStream_t s, void *work) {
    TsluWork t = tslu_carve(work);
    RLG_TRY(fill_u32(t.flag, TMAXG, 0u, s));
    return fill_u32(t.done, 1, 0u, s);
}

// Bounded wait until *p >= target (agent scope), one lane polling.
__device__ __forceinline__ void wait_ge(unsigned *p, unsigned target, int *timeout) {
    if (threadIdx.x == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            if (++spins > (1u << 22)) {
                __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_wave_barrier();
}

// Publish *p = v after every earlier store of the wave (agent scope).
__device__ __forceinline__ void publish(unsigned *p, unsigned v) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    if (threadIdx.x == 0) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// x of lane `src` (uniform) to every lane.
__device__ __forceinline__ float bcast(float x, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}

// Candidate order: larger magnitude, then the smaller row number.
__device__ __forceinline__ bool tbetter(float v2, int i2, float v1, int i1) {
    if (v2 != v1) return v2 > v1;
    return i2 < i1;
}

// Partial pivoting on up to 128 rows x w columns held by one wave: lane l
// holds row slot 0 (a0, row number id0) and slot 1 (a1, id1); negative row
// numbers mark absent rows.  On return order[j] (LDS) is the row number of
// the j-th pivot and who[j] its holder (2 * lane + slot); the rows hold
// their multipliers left of their pivot step and their U part from it on.
// zstep receives the first step whose pivot was exactly zero (or w).
__device__ __forceinline__ void tslu_gepp(float (&a0)[TW], float (&a1)[TW], int id0, int id1, int w, int *order,
                          int *who, int &zstep) {
    const int lane = threadIdx.x;
    bool p0 = id0 < 0, p1 = id1 < 0;
    zstep = w;
#pragma unroll
    for (int j = 0; j < TW; ++j) {
        if (j < w) {
            float v0 = p0 ? -1.f : fabsf(a0[j]);
            float v1 = p1 ? -1.f : fabsf(a1[j]);
            int i0 = p0 ? INT_MAX : id0, i1 = p1 ? INT_MAX : id1;
            float bv = v0;
            int bi = i0;
            if (tbetter(v1, i1, bv, bi)) { bv = v1; bi = i1; }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const float ov = __shfl_xor(bv, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (tbetter(ov, oi, bv, bi)) { bv = ov; bi = oi; }
            }
            // holder of the pivot row (uniform)
            const unsigned long long own = __ballot(bi != INT_MAX && (id0 == bi || id1 == bi));
            const int wl = own ? __ffsll((long long)own) - 1 : 0;
            const int ws = __builtin_amdgcn_readlane(id0, wl) == bi ? 0 : 1;
            if (lane == 0) { order[j] = bi == INT_MAX ? -1 : bi; who[j] = 2 * wl + ws; }
            if (lane == wl) { if (ws) p1 = true; else p0 = true; }
            const float piv = bcast(ws ? a1[j] : a0[j], wl);
            if (piv == 0.f || bi == INT_MAX) {
                zstep = min(zstep, j);
            } else {
                const float rp = 1.f / piv;
                const float m0 = p0 ? 0.f : a0[j] * rp;
                const float m1 = p1 ? 0.f : a1[j] * rp;
                if (!p0) a0[j] = m0;
                if (!p1) a1[j] = m1;
#pragma unroll
                for (int c = j + 1; c < TW; ++c) {
                    const float pc = bcast(ws ? a1[c] : a0[c], wl);
                    if (!p0) a0[c] -= m0 * pc;
                    if (!p1) a1[c] -= m1 * pc;
                }
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_tslu(float *__restrict__ A, long lda, int k, int m, int w,
                                             int64_t *__restrict__ ipiv, int64_t *__restrict__ info,
                                             int n, TsluWork tw, unsigned epoch, int *timeout) {
    __shared__ float orig[TR][TW + 1];   // this chunk's rows, original values
    __shared__ float cur[TR][TW + 1];    // the current level's candidate rows
    __shared__ float uinv[TW][TW + 1];   // wave 0: inverse of U11
    __shared__ int cur_id[TR];
    __shared__ int order[TW], who[TW];
    __shared__ int mrow[2 * TW], mpos[2 * TW];
    __shared__ int s_cnt;
    const int G = gridDim.x, g = blockIdx.x, lane = threadIdx.x;
    const unsigned base = epoch * TLEV;

    // ---- 1. load the chunk (16 rows of loads in flight before their LDS stores)
    for (int r0 = 0; r0 < TR; r0 += 16) {
        float v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = g * TR + r0 + q;
            v[q] = (r < m && lane < w) ? A[(long)(k + r) * lda + k + lane] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) orig[r0 + q][lane] = v[q];
    }
    const int id0 = g * TR + lane < m ? k + g * TR + lane : -1;
    const int id1 = g * TR + 64 + lane < m ? k + g * TR + 64 + lane : -1;
    float a0[TW], a1[TW];
    int zstep = w;
#pragma unroll
    for (int c = 0; c < TW; ++c) { a0[c] = orig[lane][c]; a1[c] = orig[64 + lane][c]; }
    tslu_gepp(a0, a1, id0, id1, w, order, who, zstep);
    __builtin_amdgcn_wave_barrier();
    // publish the winners' original rows
    for (int j = 0; j < TW; ++j) {
        // entries j >= w of order/who are never written: use row 0 for them
        const bool ok = j < w && order[j] >= 0;
        const int h = ok ? who[j] : 0, lr = (h >> 1) + 64 * (h & 1);
        tw.rows[((long)g * TW + j) * TW + lane] = ok ? orig[lr][lane] : 0.f;
        if (lane == 0) tw.ids[g * TW + j] = ok ? order[j] : -1;
    }
    publish(&tw.flag[g], base + 1);

    // ---- 2. the tree
    bool active = true;
    for (int lev = 1; (1 << (lev - 1)) < G; ++lev) {
        const int stride = 1 << lev, half = stride >> 1;
        if (g % stride != 0) { active = false; break; }
        const int p = g + half;
        if (p < G) {
            wait_ge(&tw.flag[p], base + lev, timeout);  // partner published level lev - 1
            for (int j = 0; j < TW; ++j) {
                cur[j][lane] = tw.rows[((long)g * TW + j) * TW + lane];
                cur[TW + j][lane] = tw.rows[((long)p * TW + j) * TW + lane];
            }
            cur_id[lane] = tw.ids[g * TW + lane];
            cur_id[TW + lane] = tw.ids[p * TW + lane];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int c = 0; c < TW; ++c) { a0[c] = cur[lane][c]; a1[c] = cur[64 + lane][c]; }
            tslu_gepp(a0, a1, cur_id[lane], cur_id[64 + lane], w, order, who, zstep);
            __builtin_amdgcn_wave_barrier();
            for (int j = 0; j < TW; ++j) {
                const bool ok = j < w && order[j] >= 0;
                const int h = ok ? who[j] : 0, lr = (h >> 1) + 64 * (h & 1);
                tw.rows[((long)g * TW + j) * TW + lane] = ok ? cur[lr][lane] : 0.f;
                if (lane == 0) tw.ids[g * TW + j] = ok ? order[j] : -1;
            }
        }
        publish(&tw.flag[g], base + lev + 1);
    }

    // ---- 3. wave 0: pivot block, swaps, inverse of U11
    if (g == 0 && active) {
        // a0 / a1 hold the factored rows of the last partial pivoting run
        // (the root's); row holder who[j] is pivot j
#pragma unroll
        for (int c = 0; c < TW; ++c) { cur[lane][c] = a0[c]; cur[64 + lane][c] = a1[c]; }
        __builtin_amdgcn_wave_barrier();
        for (int j = 0; j < w; ++j) {
            const int h = who[j], lr = (h >> 1) + 64 * (h & 1);
            tw.F[j * TW + lane] = cur[lr][lane];
        }
        // sequential swaps: position k + j receives pivot row order[j]
        if (lane == 0) s_cnt = w;
        for (int t = lane; t < w; t += 64) { mrow[t] = k + t; mpos[t] = k + t; }
        __builtin_amdgcn_wave_barrier();
        for (int j = 0; j < w; ++j) {
            const int want = order[j], cnt = s_cnt;
            int hit = -1;
            for (int b = 0; b < cnt && hit < 0; b += 64) {
                const int t = b + lane;
                const unsigned long long bal = __ballot(t < cnt && mrow[t] == want);
                if (bal) hit = b + __ffsll((long long)bal) - 1;
            }
            if (lane == 0) {
                if (hit < 0) { hit = cnt; mrow[hit] = want; mpos[hit] = want; s_cnt = cnt + 1; }
                ipiv[k + j] = mpos[hit];
                const int tr = mrow[j];
                mrow[j] = mrow[hit];
                mrow[hit] = tr;
            }
            __builtin_amdgcn_wave_barrier();
        }
        const int cnt = s_cnt;
        // rows whose position changes: row mrow[t] ends at mpos[t]
        for (int t = lane; t < cnt; t += 64) {
            tw.map[1 + t] = mrow[t];
            tw.map[1 + 2 * TW + t] = mpos[t];
        }
        if (lane == 0) {
            tw.map[0] = cnt;
            if (*info == (int64_t)n) {
                for (int j = 0; j < w; ++j) {
                    const float d = cur[(who[j] >> 1) + 64 * (who[j] & 1)][j];
                    if (j >= zstep || fabsf(d) < 1.1920929e-7f) { *info = k + j; break; }
                }
            }
        }
        // inverse of U11 (upper, w x w; row r of U11 is the factored row of
        // pivot r, held in cur): lane c solves U x = e_c into column c of uinv
        __builtin_amdgcn_wave_barrier();
        if (lane < w) {
            const int c = lane;
            for (int r = w - 1; r >= 0; --r) {
                const int hr = (who[r] >> 1) + 64 * (who[r] & 1);
                float sum = (r == c) ? 1.f : 0.f;
                for (int t = r + 1; t < w; ++t) sum -= cur[hr][t] * uinv[t][c];
                const float d = cur[hr][r];
                uinv[r][c] = (r > c || d == 0.f) ? 0.f : sum / d;
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (int r = 0; r < TW; ++r) tw.Uinv[r * TW + lane] = (r < w && lane < w) ? uinv[r][lane] : 0.f;
        publish(tw.done, base + TLEV - 1);
    }

    // ---- 4. L21 = A21 * inv(U11) for this chunk's non-pivot rows
    wait_ge(tw.done, base + TLEV - 1, timeout);
    for (int r = 0; r < TW; ++r) cur[r][lane] = tw.Uinv[r * TW + lane];
    const int cnt = tw.map[0];
    for (int t = lane; t < cnt; t += 64) { mrow[t] = tw.map[1 + t]; mpos[t] = tw.map[1 + 2 * TW + t]; }
    if (g == 0 && lane < w)
        for (int j = 0; j < w; ++j) A[(long)(k + j) * lda + k + lane] = tw.F[j * TW + lane];
    __builtin_amdgcn_wave_barrier();
    for (int slot = 0; slot < 2; ++slot) {
        const int lr = slot * 64 + lane;
        const int id = slot ? id1 : id0;
        if (id < 0) continue;
        // final position: from the map when the row moves, else its own
        int pos = id;
        for (int t = 0; t < cnt; ++t)
            if (mrow[t] == id) pos = mpos[t];
        if (pos < k + w) continue;  // a pivot row: written from F by wave 0
        float l[TW];
#pragma unroll
        for (int c = 0; c < TW; ++c) l[c] = 0.f;
        for (int t = 0; t < w; ++t) {
            const float at = orig[lr][t];
#pragma unroll
            for (int c = 0; c < TW; ++c) l[c] += at * cur[t][c];
        }
        float *dst = A + (long)pos * lda + k;
#pragma unroll
        for (int c = 0; c < TW; ++c)
            if (c < w) dst[c] = l[c];
    }
}

int tslu_panel(hipStream_t s, float *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
               int64_t *info, size_t n, void *work, unsigned epoch, int *timeout) {
    const unsigned G = ceil_div(m, TR);
    if (G > (unsigned)TMAXG || w > (size_t)TW || w == 0 || m < w) {
        set_error("TSLU panel of %zu x %zu is out of range", m, w);
        return RLG_E_INVALID_ARG;
    }
    RLG_PROF_W("lu_panel", s, 2.0 * (double)m * (double)w * 4.0);
    hipLaunchKernelGGL(k_tslu, dim3(G), dim3(64), 0, s, a, (long)lda, (int)k, (int)m, (int)w, ipiv, info,
                       (int)n, tslu_carve(work), epoch, timeout);
    RLG_LAUNCHED();
    return RLG_OK;
}

}  // namespace rlg
