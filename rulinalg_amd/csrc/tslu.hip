// Tournament pivoting (TSLU, the panel step of CALU: Grigori, Demmel and
// Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)) for the FP32 panels of the
// mixed-precision LU.  The plain FP64 drop-in keeps the reference's partial
// pivoting (lu.hip); the mixed factorization only has to be stable enough for
// iterative refinement, which TSLU is in practice (its growth factor behaves
// like partial pivoting's), and it needs a few reduction levels per panel
// instead of one grid-wide exchange per column.
//
// Two launches factor the m x w panel A[k:k+m, k:k+w] (w <= 64):
//
// k_tslu_tourney, one 256-thread workgroup per 256-row leaf:
//   1. each workgroup runs partial pivoting on its leaf (one row per thread,
//      in registers) and writes the row numbers of its w winners;
//   2. every tree node merges up to four children: the workgroup that
//      completes the last child (an atomic arrival count per node) loads the
//      children's winners' original rows and runs partial pivoting on those
//      256 rows; the others leave.  No workgroup ever waits for another, so
//      nothing depends on which workgroups are resident together and the
//      kernel can share the device with any other work;
//   3. the workgroup that completes the root holds, in the rows it selected,
//      the factored pivot block [L11 \ U11] (partial pivoting on original
//      rows); it writes the block, U11^-1, L11^-1 (for the block-row solve),
//      the reference's sequential swap list ipiv[k + j] and the list of rows
//      that the swaps move.
// k_tslu_apply, one workgroup per 64 panel rows: L21 = A21 U11^-1 on the
//   matrix cores for the rows that are not pivots, each written to its final
//   position; the first workgroup writes the pivot block to rows k..k+w-1.
#include <climits>

#include "kernels.hpp"
#include "mfma.hpp"
#include "wave.hpp"

namespace rlg {

constexpr int TW = 64;    // panel width
constexpr int TLR = 256;  // rows per leaf = threads per workgroup
constexpr int TF = 4;     // children per tree node: 4 x 64 candidates = 256 rows
static_assert(TF * TW == TLR, "a merge has as many rows as a leaf");

// Tree of L leaves: level 0 has L nodes, level l + 1 has ceil(n_l / TF).
static size_t tree_nodes(size_t leaves) {
    size_t tot = 0, c = leaves;
    for (;;) {
        tot += c;
        if (c == 1) return tot;
        c = (c + TF - 1) / TF;
    }
}

struct TsluWork {
    int *cand;        // [nodes][TW] row numbers of a node's winners (-1: none)
    unsigned *cnt;    // [nodes] arrival counts, zero between panels
    float *F;         // [TW][TW] factored pivot block, row j = pivot j
    float *Uinv;      // [TW][TW] U11^-1, zero outside w x w
    int *sel;         // [TW] pivot rows in order
    int *map;         // [1 + 2 TW]: count, then (source row, destination) pairs
    int leaves_max;
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t leaves_for(size_t rows) { return (rows + TLR - 1) / TLR; }

size_t tslu_work_bytes(size_t n) {
    const size_t nodes = tree_nodes(leaves_for(n ? n : 1));
    return al(nodes * TW * 4) + al(nodes * 4) + 2 * al(TW * TW * 4) + al(TW * 4) + al((1 + 2 * TW) * 4);
}

static TsluWork tslu_carve(void *work, size_t n) {
    const size_t nodes = tree_nodes(leaves_for(n ? n : 1));
    char *p = static_cast<char *>(work);
    TsluWork t;
    t.cand = reinterpret_cast<int *>(p); p += al(nodes * TW * 4);
    t.cnt = reinterpret_cast<unsigned *>(p); p += al(nodes * 4);
    t.F = reinterpret_cast<float *>(p); p += al(TW * TW * 4);
    t.Uinv = reinterpret_cast<float *>(p); p += al(TW * TW * 4);
    t.sel = reinterpret_cast<int *>(p); p += al(TW * 4);
    t.map = reinterpret_cast<int *>(p);
    t.leaves_max = (int)leaves_for(n ? n : 1);
    return t;
}

int tslu_reset(hipStream_t s, void *work, size_t n) {
    const size_t nodes = tree_nodes(leaves_for(n ? n : 1));
    return fill_u32(tslu_carve(work, n).cnt, nodes, 0u, s);
}

// Maximum of a 32-bit key over the wave (uniform): two quad swaps and two row
// mirrors, then the four rows' maxima.
__device__ __forceinline__ unsigned wave_max_u32(unsigned k) {
    k = max(k, (unsigned)dpp<DPP_XOR1>((int)k));
    k = max(k, (unsigned)dpp<DPP_XOR2>((int)k));
    k = max(k, (unsigned)dpp<DPP_HALF_MIRROR>((int)k));
    k = max(k, (unsigned)dpp<DPP_MIRROR>((int)k));
    const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)k, 0);
    const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)k, 16);
    const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)k, 32);
    const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)k, 48);
    return max(max(a, b), max(c, d));
}

// Shared state of one partial-pivoting run over 256 rows.
struct GeppShared {
    float cand[2][4][TW];   // each wave's candidate row, by column parity
    unsigned key[2][4];     // its |value| bits + 1 (0: the wave has no row left)
    int id[2][4];           // its row number
    int order[TW];          // row number of pivot j (-1: none)
};

// Partial pivoting on the 256 rows held one per thread (a[] = the row, id =
// its row number or -1): at column j every wave proposes its largest
// |a[j]| among the rows not yet chosen, the largest of the four proposals
// is pivot j (ties: the lower wave, then the lower lane; the pivot choice is
// free for the tournament), and the other rows are eliminated with it.  On
// return the chosen row of step j holds its multipliers left of column j and
// its U row from column j on; `step` is that j for the chosen rows, -1 for
// the others.  One workgroup barrier per column.
__device__ __forceinline__ void gepp256(float (&a)[TW], int id, int w, GeppShared &sh, int &step) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    step = -1;
    static_for<TW>([&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if (j < w) {
            constexpr int par = j & 1;
            const bool active = id >= 0 && step < 0;
            const unsigned key = active ? __float_as_uint(fabsf(a[j])) + 1u : 0u;
            const unsigned wmax = wave_max_u32(key);
            if (wmax != 0u) {
                const unsigned long long hit = __ballot(key == wmax);
                const int wl = __ffsll((long long)hit) - 1;
                if (lane == wl) {
                    constexpr int c0 = j & ~3;
#pragma unroll
                    for (int c = c0; c < TW; c += 4)
                        *reinterpret_cast<float4 *>(&sh.cand[par][wave][c]) =
                            make_float4(a[c], a[c + 1], a[c + 2], a[c + 3]);
                    sh.id[par][wave] = id;
                }
            }
            if (lane == 0) sh.key[par][wave] = wmax;
            __syncthreads();
            unsigned best = sh.key[par][0];
            int ww = 0;
#pragma unroll
            for (int q = 1; q < 4; ++q) {
                const unsigned kq = sh.key[par][q];
                if (kq > best) { best = kq; ww = q; }
            }
            const int pid = best ? sh.id[par][ww] : -1;
            if (tid == 0) sh.order[j] = pid;
            if (best && id == pid) {
                step = j;
            } else if (best && active) {
                const float p = sh.cand[par][ww][j];
                const float m = p != 0.f ? a[j] * __builtin_amdgcn_rcpf(p) : 0.f;
                a[j] = m;
#pragma unroll
                for (int c = j + 1; c < TW; ++c) a[c] = fmaf(-m, sh.cand[par][ww][c], a[c]);
            }
        } else {
            if (tid == 0) sh.order[j] = -1;
        }
    });
    __syncthreads();
}

// Row `id` of the panel (columns k .. k+w-1, zeros beyond) into a[].
__device__ __forceinline__ void load_row(float (&a)[TW], const float *__restrict__ A, long lda, int k,
                                         int id, int w, bool vec) {
    if (id < 0) {
#pragma unroll
        for (int c = 0; c < TW; ++c) a[c] = 0.f;
        return;
    }
    const float *r = A + (long)id * lda + k;
    if (vec && w == TW) {
#pragma unroll
        for (int c = 0; c < TW; c += 4) {
            const float4 v = *reinterpret_cast<const float4 *>(r + c);
            a[c] = v.x; a[c + 1] = v.y; a[c + 2] = v.z; a[c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < TW; ++c) a[c] = c < w ? r[c] : 0.f;
    }
}

// In-place inverses of two lower-triangular 64 x 64 matrices, M0 with a unit
// diagonal and M1 with its own (the upper parts are zero and stay zero).
// Each is split in 32 x 32 halves: the two diagonal halves of M0 are
// inverted by waves 0 and 1 and those of M1 by waves 2 and 3, side by side,
// lane j forward-substituting column j of the identity with the column in
// registers; the corners are X21 = -X22 (M21 X11), two 32 x 32 products on
// the matrix cores per matrix, one 16 x 16 tile per wave.  `tmp` is scratch.
// Called by the whole workgroup.
__device__ void inv_lower64_pair(float (*M0)[TW + 1], float (*M1)[TW + 1], float (*tmp)[32][33]) {
    using Acc = Mfma<float>::acc;
    constexpr int H = TW / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    if (lane < H) {
        float (*M)[TW + 1] = wave < 2 ? M0 : M1;
        const bool unit = wave < 2;
        const int o = (wave & 1) * H, j = lane;
        float x[H];
#pragma unroll
        for (int i = 0; i < H; ++i) {
            float s = (i == j) ? 1.f : 0.f;
#pragma unroll
            for (int q = 0; q < i; ++q) s -= M[o + i][o + q] * x[q];
            x[i] = unit ? s : s / M[o + i][o + i];
        }
        // every read of the half precedes the wave's writes to it
#pragma unroll
        for (int i = 0; i < H; ++i) M[o + i][o + j] = x[i];
    }
    __syncthreads();
    const int tr = (wave >> 1) * 16, tc = (wave & 1) * 16;
#pragma unroll
    for (int q = 0; q < 2; ++q) {   // tmp = M21 X11
        float (*M)[TW + 1] = q ? M1 : M0;
        Acc acc = Acc{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < H / 4; ++s) acc = Mfma<float>::op(M[H + tr + lr][4 * s + lk], M[4 * s + lk][tc + lr], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) tmp[q][tr + Mfma<float>::row(lk, r)][tc + lr] = acc[r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {   // X21 = -X22 tmp, over M21
        float (*M)[TW + 1] = q ? M1 : M0;
        Acc acc = Acc{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < H / 4; ++s)
            acc = Mfma<float>::op(M[H + tr + lr][H + 4 * s + lk], tmp[q][4 * s + lk][tc + lr], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) M[H + tr + Mfma<float>::row(lk, r)][tc + lr] = -acc[r];
    }
    __syncthreads();
}

// Shared memory of the tournament kernel: the partial-pivoting state, and for
// the root the pivot block and the two triangles being inverted.
struct TourneyShared {
    GeppShared g;
    float F[TW][TW + 1];
    float L[TW][TW + 1];   // L11, then L11^-1
    float Ut[TW][TW + 1];  // U11^T, then (U11^-1)^T
    float tmp[2][32][33];
    int slot_sel[TW];      // top slot q -> pivot index whose row starts there (-1)
    int vrow[TW], vsel[TW];
    int s_last;
};

__global__ __launch_bounds__(TLR) void k_tslu_tourney(const float *__restrict__ A, long lda, int k, int m,
                                                      int w, int64_t *__restrict__ ipiv,
                                                      int64_t *__restrict__ info, int n,
                                                      float *__restrict__ linv, TsluWork tw, int vec) {
    __shared__ TourneyShared sh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = gridDim.x;
    float a[TW];
    int step = -1;

    // ---- leaf
    const int r0 = blockIdx.x * TLR;
    int id = r0 + tid < m ? k + r0 + tid : -1;
    load_row(a, A, lda, k, id, w, vec != 0);
    gepp256(a, id, w, sh.g, step);
    int level_off = 0, level_n = L, node = blockIdx.x;
    if (tid < TW) tw.cand[(level_off + node) * TW + tid] = sh.g.order[tid];

    // ---- merges: the workgroup completing a node's last child merges it
    while (level_n > 1) {
        const int parent = node / TF;
        const int next_off = level_off + level_n, next_n = (level_n + TF - 1) / TF;
        const int kids = min(TF, level_n - parent * TF);
        __threadfence();  // this workgroup's winners before its arrival
        __syncthreads();
        if (tid == 0) {
            unsigned *c = &tw.cnt[next_off + parent];
            const unsigned old = atomicAdd(c, 1u);
            sh.s_last = old == (unsigned)(kids - 1);
            if (sh.s_last) *c = 0u;  // every child has arrived: ready for the next panel
        }
        __syncthreads();
        if (!sh.s_last) return;
        __threadfence();  // the other children's winners after their arrivals
        const int child = parent * TF + tid / TW;
        id = tid / TW < kids ? __hip_atomic_load(&tw.cand[(level_off + child) * TW + (tid % TW)],
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : -1;
        load_row(a, A, lda, k, id, w, vec != 0);
        gepp256(a, id, w, sh.g, step);
        level_off = next_off;
        level_n = next_n;
        node = parent;
        if (level_n > 1 && tid < TW) tw.cand[(level_off + node) * TW + tid] = sh.g.order[tid];
    }

    // ---- root: the pivot block, its inverses, the swaps
    if (step >= 0) {
#pragma unroll
        for (int c = 0; c < TW; ++c) sh.F[step][c] = a[c];
    }
    if (tid < TW) {
        sh.slot_sel[tid] = -1;
        if (tid >= w)
#pragma unroll
            for (int c = 0; c < TW; ++c) sh.F[tid][c] = (c == tid) ? 1.f : 0.f;
    }
    __syncthreads();
    // L11 (unit lower) and U11^T, zero-padded to 64 x 64 with unit diagonal
    for (int e = tid; e < TW * TW; e += TLR) {
        const int r = e / TW, c = e % TW;
        const float f = sh.F[r][c];
        sh.L[r][c] = c < r ? (r < w && c < w ? f : 0.f) : 0.f;
        sh.Ut[c][r] = c >= r ? (r < w && c < w ? f : (r == c ? 1.f : 0.f)) : 0.f;
        tw.F[e] = f;
    }
    __syncthreads();
    inv_lower64_pair(sh.L, sh.Ut, sh.tmp);
    for (int e = tid; e < TW * TW; e += TLR) {
        const int r = e / TW, c = e % TW;
        const bool in = r < w && c < w;
        linv[e] = in && c <= r ? sh.L[r][c] : 0.f;
        tw.Uinv[e] = in && c >= r ? sh.Ut[c][r] : 0.f;
    }
    if (wave == 0) {
        // pivot j's row: in the top slot q = row - k when it starts in
        // k .. k+w-1, else in the next free "outside" slot (half 1)
        const int selj = lane < w ? sh.g.order[lane] : -1;
        if (lane < w) tw.sel[lane] = selj;
        const bool top = selj >= k && selj < k + w;
        if (lane < w && top) sh.slot_sel[selj - k] = lane;
        // (a pivot without a row cannot occur when m >= w; it would leave
        // its step without a swap, and info reports it)
        const bool out = lane < w && selj >= 0 && !top;
        const unsigned long long bo = __ballot(out);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bo >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)bo, 0));
        if (out) { sh.vrow[rank] = selj; sh.vsel[rank] = lane; }
        const int nout = __popcll(bo);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        // slot (lane, half 0) = position k + lane; (lane, half 1) = position vrow[lane]
        int pos0 = k + lane, con0 = k + lane, sx0 = lane < w ? sh.slot_sel[lane] : -1;
        int pos1 = lane < nout ? sh.vrow[lane] : -1, con1 = pos1, sx1 = lane < nout ? sh.vsel[lane] : -1;
        // the reference's sequential swaps: step j exchanges position k + j
        // with the position that holds pivot j's row now
        for (int j = 0; j < w; ++j) {
            const unsigned long long b0 = __ballot(sx0 == j), b1 = __ballot(sx1 == j);
            if (!b0 && !b1) {
                if (lane == 0) ipiv[k + j] = k + j;
                continue;
            }
            const int H = b0 ? 0 : 1;
            const int Ln = __ffsll((long long)(b0 ? b0 : b1)) - 1;
            const int psl = __builtin_amdgcn_readlane(H ? pos1 : pos0, Ln);
            const int csl = __builtin_amdgcn_readlane(H ? con1 : con0, Ln);
            const int ctop = __builtin_amdgcn_readlane(con0, j), xtop = __builtin_amdgcn_readlane(sx0, j);
            if (lane == 0) ipiv[k + j] = psl;
            if (!(H == 0 && Ln == j)) {
                if (lane == j) { con0 = csl; sx0 = j; }
                if (lane == Ln) {
                    if (H) { con1 = ctop; sx1 = xtop; }
                    else { con0 = ctop; sx0 = xtop; }
                }
            }
        }
        // rows that end outside the top: the displaced rows
        const bool mv = lane < nout && con1 != pos1;
        const unsigned long long bm = __ballot(mv);
        const int at = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0));
        if (mv) { tw.map[1 + 2 * at] = con1; tw.map[2 + 2 * at] = pos1; }
        if (lane == 0) {
            tw.map[0] = __popcll(bm);
            if (*info == (int64_t)n) {
                for (int j = 0; j < w; ++j) {
                    const float d = sh.F[j][j];
                    if (sh.g.order[j] < 0 || !(fabsf(d) >= 1.1920929e-7f)) { *info = k + j; break; }
                }
            }
        }
    }
}

// L21 = A21 U11^-1 for panel rows [k + 64 b, k + 64 b + 64), each written to
// its final position; the pivot rows are skipped (their factored rows come
// from the pivot block, which workgroup 0 writes over rows k .. k+w-1 after
// reading its own rows).  The destinations of moved rows are positions of
// pivot rows, whose original values no workgroup needs, so the in-place
// writes cannot race with another workgroup's reads.
__global__ __launch_bounds__(256) void k_tslu_apply(float *__restrict__ A, long lda, int k, int m, int w,
                                                    TsluWork tw) {
    using Acc = Mfma<float>::acc;
    __shared__ float Ui[TW][TW + 1];
    __shared__ float X[64][TW + 1];
    __shared__ int dst[64];
    __shared__ unsigned long long selmask;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int base = k + (int)blockIdx.x * 64;
    const int rows = min(64, k + m - base);
    {
        constexpr int PER = 64 * TW / 256;
        float vu[PER], vx[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e / TW, c = e % TW;
            vu[q] = tw.Uinv[e];
            vx[q] = (r < rows && c < w) ? A[(long)(base + r) * lda + k + c] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = q * 256 + tid, r = e / TW, c = e % TW;
            Ui[r][c] = vu[q];
            X[r][c] = vx[q];
        }
    }
    if (tid < 64) dst[tid] = base + tid;
    if (tid == 0) selmask = 0ull;
    __syncthreads();
    if (tid < w) {
        const int s = tw.sel[tid];
        if (s >= base && s < base + 64) atomicOr(&selmask, 1ull << (s - base));
    }
    {
        const int cnt = tw.map[0];
        if (tid < cnt) {
            const int src = tw.map[1 + 2 * tid], d = tw.map[2 + 2 * tid];
            if (src >= base && src < base + 64) dst[src - base] = d;
        }
    }
    __syncthreads();
    Acc acc[4];
#pragma unroll
    for (int jc = 0; jc < 4; ++jc) acc[jc] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < TW / 4; ++s) {
        const float av = X[16 * wave + lr][4 * s + lk];
#pragma unroll
        for (int jc = 0; jc < 4; ++jc) acc[jc] = Mfma<float>::op(av, Ui[4 * s + lk][16 * jc + lr], acc[jc]);
    }
    const unsigned long long sm = selmask;
#pragma unroll
    for (int jc = 0; jc < 4; ++jc)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * wave + Mfma<float>::row(lk, r), col = 16 * jc + lr;
            if (row < rows && col < w && !((sm >> row) & 1ull))
                A[(long)dst[row] * lda + k + col] = acc[jc][r];
        }
    if (blockIdx.x == 0) {
        for (int e = tid; e < w * TW; e += 256) {
            const int r = e / TW, c = e % TW;
            if (c < w) A[(long)(k + r) * lda + k + c] = tw.F[e];
        }
    }
}

size_t tslu_max_rows(size_t n) { return leaves_for(n ? n : 1) * TLR; }

int tslu_panel(hipStream_t s, float *a, size_t lda, size_t k, size_t m, size_t w, int64_t *ipiv,
               int64_t *info, size_t n, void *work, float *linv) {
    const TsluWork tw = tslu_carve(work, n);
    const size_t L = leaves_for(m);
    if (L > (size_t)tw.leaves_max || w > (size_t)TW || w == 0 || m < w) {
        set_error("TSLU panel of %zu x %zu is out of range", m, w);
        return RLG_E_INVALID_ARG;
    }
    // row loads are 16-byte vectors when every panel row starts 16-byte aligned
    const int vec = (lda % 4 == 0 && k % 4 == 0 && ((uintptr_t)a & 15) == 0) ? 1 : 0;
    {
        // algorithmic bytes: the m x w panel read once (the leaves)
        RLG_PROF_W("tslu_tourney", s, (double)m * (double)w * 4.0);
        hipLaunchKernelGGL(k_tslu_tourney, dim3((unsigned)L), dim3(TLR), 0, s, a, (long)lda, (int)k, (int)m,
                           (int)w, ipiv, info, (int)n, linv, tw, vec);
        RLG_LAUNCHED();
    }
    {
        // algorithmic bytes: the panel read and written once
        RLG_PROF_W("tslu_apply", s, 2.0 * (double)m * (double)w * 4.0);
        hipLaunchKernelGGL(k_tslu_apply, dim3(ceil_div(m, 64)), dim3(256), 0, s, a, (long)lda, (int)k,
                           (int)m, (int)w, tw);
        RLG_LAUNCHED();
    }
    return RLG_OK;
}

}  // namespace rlg
