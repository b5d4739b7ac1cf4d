// Batched factor-and-solve of many independent systems of uneven order in
// one persistent launch.
//
// Grid: one workgroup per CU (the kernel's registers allow no second); each
// workgroup takes systems from a device work counter (longest first, the
// order the host computes) until the list is exhausted, so every workgroup
// reaches the exit test and the grid drains.  One workgroup factors and solves one
// system entirely:
//
//   LU (PartialPivLu::decompose + solve, lu.rs:163-244): right-looking,
//   blocked by NB columns.  The (n - k) x NB panel lives in LDS and is
//   factored with the reference's pivot rule (first maximum of |a|, strict
//   `>`), division multipliers and separately rounded updates; its swaps are
//   applied to the rest of the matrix and to the right-hand sides; the block
//   row is solved in registers; the trailing matrix is updated with
//   v_mfma_f64_16x16x4f64 tiles whose A-fragments come from the LDS panel.
//   The right-hand sides ride along as extra columns (forward substitution
//   fused into the elimination), then a blocked back substitution finishes.
//
//   Cholesky (cholesky.rs:116-202): the same skeleton with the reference's
//   left-looking recurrence inside each diagonal block, lower triangle only,
//   the SYRK update on lower tiles, and forward / transposed-back
//   substitutions.
//
// NB (8, 16 or 32) is chosen per system from an autotuned table of n-buckets
// (rlg_batch_autotune); (n + 16) * (NB + 1) doubles must fit the panel
// buffer.  The mixed form factors an FP32 copy in a per-workgroup workspace
// and refines in FP64 against the untouched FP64 matrix.
//
// This header holds the device code only: the kernels and the phases they
// call.  It depends on wave.hpp, on the item type and status codes of
// include/rlgpu.h and on a small set of device facilities (thread and block
// indices, workgroup barriers, lane reads, DPP moves, ballots, lane counts,
// the FP64 and FP32 16x16x4 MFMA, atomicAdd), so that a host build that
// supplies those runs the same source on the CPU.
#pragma once

#include <climits>
#include <cmath>
#include <cstdint>

#include "../../include/rlgpu.h"
#include "wave.hpp"

namespace rlg {

constexpr int BTH = 256;
constexpr int BMAXN = RLG_BATCH_MAX_N;
constexpr int PANEL = 8704;       // doubles: 512 x 17
constexpr int BMAXR = RLG_BATCH_MAX_NRHS;

typedef double bd4 __attribute__((ext_vector_type(4)));
typedef float bf4 __attribute__((ext_vector_type(4)));

template <class T> struct BMfma;
template <> struct BMfma<double> {
    using acc = bd4;
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct BMfma<float> {
    using acc = bf4;
    static __device__ __forceinline__ int row(int lk, int r) { return lk * 4 + r; }
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <class T> __device__ __forceinline__ T beps();
template <> __device__ __forceinline__ double beps<double>() { return 2.220446049250313e-16; }
template <> __device__ __forceinline__ float beps<float>() { return 1.1920929e-7f; }

struct Smem {
    double panel[PANEL];
    double red_v[BTH / 64];
    int red_i[BTH / 64];
    int piv[32];
    double div[32];
    // LU column step j (buffers by the parity of j): each wave's pivot
    // candidate (|value|, row, the whole row) and row j of the panel
    double cand_v[2][BTH / 64];
    int cand_i[2][BTH / 64];
    double cand[2][BTH / 64][32];
    double jrow[2][32];
    int mdst[64], msrc[64];     // net row moves of a panel's swaps
    int mcnt;
    int info;
    int item;
};

constexpr int BNW = BTH / 64;  // waves per workgroup
constexpr int TSTRIP = 4;      // 16 x 16 tiles per wave step of the Cholesky SYRK update
constexpr int TSTRIP_LU = 8;   // and of the LU trailing update (measured 15 % faster than 4;
                               // the SYRK's strips end at the diagonal, where 8 wastes tiles)

// Copy the m x kb block at A[k.., k..] (ld n) into the LDS panel P (ld LD), or
// back (TO_A); eight loads in flight per thread before their stores.  With
// LOWER only the lower triangle of the top kb x kb block is moved.
template <class T, int LD, bool TO_A, bool LOWER>
__device__ __forceinline__ void panel_move(T *__restrict__ A, int n, int k, int m, int kb,
                                           T *__restrict__ P) {
    const int tot = m * kb, tid = threadIdx.x;
    for (int e0 = tid; e0 < tot; e0 += 8 * BTH) {
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = e0 + q * BTH, r = e / kb, c = e - r * kb;
            const bool ok = e < tot && (!LOWER || r >= kb || c <= r);
            if (TO_A) v[q] = ok ? P[r * LD + c] : T(0);
            else v[q] = ok ? A[(long)(k + r) * n + k + c] : T(0);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = e0 + q * BTH, r = e / kb, c = e - r * kb;
            const bool ok = e < tot && (!LOWER || r >= kb || c <= r);
            if (ok) {
                if (TO_A) A[(long)(k + r) * n + k + c] = v[q];
                else P[r * LD + c] = v[q];
            }
        }
    }
}

// Net effect of a panel's swaps (swap s exchanges local rows s and piv[s] >=
// s, s < kb) as moves (dst, src): afterwards row dst holds what row src held
// before.  Row t < kb is final after swap t and holds what row piv[t] held
// just before it; a row x beyond the block is final after the last swap that
// names it, and then holds what row s held just before swap s.  "What row x
// held before swap t" is found by scanning the earlier swaps downwards.
// Called by every thread; wave 0 (kb <= 32 lanes) fills sm.mdst / sm.msrc.
__device__ __forceinline__ int bchase(const int *piv, int x, int t) {
    for (int s = t - 1; s >= 0; --s)
        if (piv[s] == x) x = s;
    return x;
}
__device__ __forceinline__ void panel_moves(int kb, const int *piv, int *mdst, int *msrc, int *mcnt) {
    const int t = threadIdx.x;
    if (t < 64) {
        int d_in = -1, s_in = -1, d_out = -1, s_out = -1;
        if (t < kb) {
            const int pt = piv[t];
            const int si = bchase(piv, pt, t);
            if (si != t) { d_in = t; s_in = si; }
            if (pt >= kb) {
                bool last = true;
                for (int s = t + 1; s < kb; ++s) last = last && piv[s] != pt;
                if (last) { d_out = pt; s_out = bchase(piv, t, t); }
            }
        }
        const unsigned long long b_in = __ballot(d_in >= 0), b_out = __ballot(d_out >= 0);
        const int n_in = __popcll(b_in);
        const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b_in >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b_in, 0));
        const int below_o = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b_out >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b_out, 0));
        if (d_in >= 0) { mdst[below] = d_in; msrc[below] = s_in; }
        if (d_out >= 0) { mdst[n_in + below_o] = d_out; msrc[n_in + below_o] = s_out; }
        if (t == 0) *mcnt = n_in + __popcll(b_out);
    }
}

// ---------------------------------------------------------------- LU

// Factor the n x n matrix A (ld = n, type T) in place with RHS vectors B
// (nrhs x n, may be null) carried along as extra columns, so that on return
// B holds L^-1 P B.  Pivots go to piv_out[0..n).  Returns the first failing
// step or n.
template <class T, int NB>
__device__ int lu_fused(T *__restrict__ A, int n, T *__restrict__ B, int nrhs,
                        int64_t *__restrict__ piv_out, Smem &sm) {
    T *P = reinterpret_cast<T *>(sm.panel);
    constexpr int LD = NB + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) sm.info = n;
    __syncthreads();
    for (int k = 0; k < n; k += NB) {
        const int kb = min(NB, n - k), m = n - k;
        panel_move<T, LD, false, false>(A, n, k, m, kb, P);
        __syncthreads();
        // Column j of the panel, one barrier: before it, each wave's best
        // candidate for column j (|value|, row and the whole row) and row j
        // are copied to buffers of the column's parity; after it every
        // thread picks the same pivot row p from the four candidates, threads
        // 0..kb-1 write the pivot row into row j, and each thread eliminates
        // its own rows, taking the pivot row and (for row p) row j from the
        // copies, then searches column j + 1 over its rows and the waves
        // publish the next candidates into the other parity's buffers.
        // Thread t owns the panel rows t + q BTH and is the only one to write
        // them, apart from row j.  Rows are read across threads only through
        // the copies, which a wave makes from rows its own lanes wrote (for
        // column 0, from the loaded panel, after the barrier that follows
        // the load).
        constexpr int RQ = NB == 32 ? 1 : 2;  // NB = 32 only for m <= 247
        T *cand = reinterpret_cast<T *>(&sm.cand[0][0][0]);  // [2][BNW][NB]
        T *jrows = reinterpret_cast<T *>(&sm.jrow[0][0]);     // [2][NB]
        // publish this wave's candidate (bv, bi: the lane's own) for the
        // column of parity `par`, and (wave 0) row `jn` set aside
        auto publish = [&](T bv, int bi, int par, int jn) __attribute__((always_inline)) {
            const int wl = wave_first_max_lane(bv, bi);
            bv = wl >= 0 ? lane_get(bv, wl) : T(0);
            bi = wl >= 0 ? lane_get(bi, wl) : INT_MAX;
            if (lane < NB)
                cand[(par * BNW + wave) * NB + lane] = (bi != INT_MAX && lane < kb) ? P[bi * LD + lane] : T(0);
            if (lane == 0) { sm.cand_v[par][wave] = (double)bv; sm.cand_i[par][wave] = bi; }
            if (wave == 0 && lane < NB) jrows[par * NB + lane] = lane < kb ? P[jn * LD + lane] : T(0);
        };
        {
            T bv = T(0);
            int bi = INT_MAX;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int r = tid + q * BTH;
                const T v = r < m ? fabs(P[r * LD]) : T(0);
                const bool take = (r < m) & better(v, r, bv, bi);
                bv = take ? v : bv;
                bi = take ? r : bi;
            }
            publish(bv, bi, 0, 0);
        }
        __syncthreads();
        for (int j = 0; j < kb; ++j) {
#pragma clang fp contract(off)
            const int par = j & 1;
            T pa = (T)sm.cand_v[par][0];
            int p = sm.cand_i[par][0], ww = 0;
#pragma unroll
            for (int q = 1; q < BNW; ++q) {
                const bool take = better((T)sm.cand_v[par][q], sm.cand_i[par][q], pa, p);
                pa = take ? (T)sm.cand_v[par][q] : pa;
                p = take ? sm.cand_i[par][q] : p;
                ww = take ? q : ww;
            }
            p = __builtin_amdgcn_readfirstlane(p);  // the same in every lane
            ww = __builtin_amdgcn_readfirstlane(ww);
            const T *prow = cand + (par * BNW + ww) * NB;  // row p (zeros at columns >= kb)
            const T *jrow = jrows + par * NB;              // row j (zeros at columns >= kb)
            const T piv = prow[j];
            if (tid == 0) {
                sm.piv[j] = p;
                if (pa < beps<T>() && sm.info == n) sm.info = k + j;
            }
            if (tid < kb) P[j * LD + tid] = prow[tid];
            T bv = T(0);
            int bi = INT_MAX;
#pragma unroll
            for (int q = 0; q < RQ; ++q) {
                const int r = tid + q * BTH;
                if (r < m && r > j) {
                    const T *src = r == p ? jrow : P + r * LD;  // row r after the swap
                    const T mult = src[j] / piv;
                    T x[NB], pm[NB];
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        x[c] = src[c];
                        pm[c] = c > j ? prow[c] : T(0);
                    }
#pragma unroll
                    for (int c = 0; c < NB; ++c) P[r * LD + c] = x[c] - mult * pm[c];
                    P[r * LD + j] = mult;
                    if (j + 1 < kb) {
                        const T v = fabs(P[r * LD + j + 1]);
                        const bool take = better(v, r, bv, bi);
                        bv = take ? v : bv;
                        bi = take ? r : bi;
                    }
                }
            }
            if (j + 1 < kb) publish(bv, bi, par ^ 1, j + 1);
            __syncthreads();
        }
        if (tid < kb) piv_out[k + tid] = k + sm.piv[tid];
        panel_move<T, LD, true, false>(A, n, k, m, kb, P);
        // row swaps on the columns outside the panel and on the RHS, as the
        // net moves of the panel's swaps: every source read before any write
        panel_moves(kb, sm.piv, sm.mdst, sm.msrc, &sm.mcnt);
        __syncthreads();
        const int right = n - k - kb;
        const int ncols = k + right + nrhs;
        const int mc = __builtin_amdgcn_readfirstlane(sm.mcnt);
        if (mc > 0) {
            for (int q = tid; q < ncols; q += BTH) {
                T *base;
                long rs;
                if (q < k) { base = A + (long)k * n + q; rs = n; }
                else if (q < k + right) { base = A + (long)k * n + kb + q; rs = n; }
                else { base = B + (long)(q - k - right) * n + k; rs = 1; }
                T v[2 * NB];
#pragma unroll
                for (int e = 0; e < 2 * NB; ++e)
                    if (e < mc) v[e] = base[sm.msrc[e] * rs];
#pragma unroll
                for (int e = 0; e < 2 * NB; ++e)
                    if (e < mc) base[sm.mdst[e] * rs] = v[e];
            }
        }
        __syncthreads();
        // block row: U12 = L11^-1 A12 and Y1 = L11^-1 B1, one column per thread,
        // column sweep as gaussian_elimination applies it (lu.rs:603-616)
        for (int q = tid; q < right + nrhs; q += BTH) {
            T u[NB];
            T *base;
            long stride;
            if (q < right) { base = A + (long)k * n + k + kb + q; stride = n; }
            else { base = B + (long)(q - right) * n + k; stride = 1; }
#pragma unroll
            for (int i = 0; i < NB; ++i) u[i] = i < kb ? base[i * stride] : T(0);
            {
#pragma clang fp contract(off)
                // unguarded: in a last panel narrower than NB the entries
                // u[kb..] take garbage from LDS rows past the panel and are
                // never stored
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    // one column of L11 at a time: scheduled freely, the
                    // unrolled sweep would hold all of L11 in registers
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int i = c + 1; i < NB; ++i) u[i] = u[i] - P[i * LD + c] * u[c];
                }
            }
#pragma unroll
            for (int i = 1; i < NB; ++i)
                if (i < kb) base[i * stride] = u[i];
        }
        __syncthreads();
        // trailing update A22 -= L21 * U12 on MFMA tiles: a wave takes a strip
        // of TSTRIP_LU tiles along a tile row and issues all of the strip's
        // global loads (C tiles and U12 fragments) before its MFMAs
        const int mt = m - kb;  // rows / cols of A22
        if (mt > 0) {
            using Acc = typename BMfma<T>::acc;
            constexpr int KS = NB / 4;
            const int tiles = (mt + 15) / 16, strips = (tiles + TSTRIP_LU - 1) / TSTRIP_LU;
            const int lr = lane & 15, lk = lane >> 4;
            T *C = A + (long)(k + kb) * n + k + kb;
            const T *U = A + (long)k * n + k + kb;
            for (int t = wave; t < tiles * strips; t += BNW) {
                const int ti = t / strips, tj0 = (t - ti * strips) * TSTRIP_LU;
                Acc acc[TSTRIP_LU];
                T bf[TSTRIP_LU][KS];
#pragma unroll
                for (int s = 0; s < TSTRIP_LU; ++s) {
                    const int gj = (tj0 + s) * 16 + lr;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int gi = ti * 16 + BMfma<T>::row(lk, r);
                        acc[s][r] = (gi < mt && gj < mt) ? C[(long)gi * n + gj] : T(0);
                    }
#pragma unroll
                    for (int q = 0; q < KS; ++q) {
                        const int kk = 4 * q + lk;
                        bf[s][q] = (gj < mt && kk < kb) ? U[(long)kk * n + gj] : T(0);
                    }
                }
                T af[KS];
#pragma unroll
                for (int q = 0; q < KS; ++q) {
                    const int kk = 4 * q + lk, ai = ti * 16 + lr;
                    af[q] = (ai < mt && kk < kb) ? -P[(kb + ai) * LD + kk] : T(0);
                }
#pragma unroll
                for (int s = 0; s < TSTRIP_LU; ++s)
#pragma unroll
                    for (int q = 0; q < KS; ++q) acc[s] = BMfma<T>::op(af[q], bf[s][q], acc[s]);
#pragma unroll
                for (int s = 0; s < TSTRIP_LU; ++s) {
                    const int gj = (tj0 + s) * 16 + lr;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int gi = ti * 16 + BMfma<T>::row(lk, r);
                        if (gi < mt && gj < mt) C[(long)gi * n + gj] = acc[s][r];
                    }
                }
            }
            // RHS rows below the block: B[r][k+kb+i] -= sum_c L21[i][c] * Y1[r][c]
            for (int q = tid; q < mt * nrhs; q += BTH) {
                const int r = q / mt, i = q - r * mt;
                T *br = B + (long)r * n;
                T s = br[k + kb + i];
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < kb) s = s - P[(kb + i) * LD + c] * br[k + c];
                br[k + kb + i] = s;
            }
        }
        __syncthreads();
    }
    return sm.info;
}

enum { TS_LOWER_UNIT = 0, TS_LOWER = 1, TS_UPPER = 2, TS_LOWER_T = 3 };

// In-place triangular solve on nrhs vectors X (vector r at X + r*n) with the
// triangle of A named by MODE, blocked by 32 rows.  The diagonal block and the
// block of X are staged in LDS; each wave sweeps two vectors of the block in
// registers (half-wave per vector, lane i holding x[i], x[c] broadcast by a
// lane read), so the sweep has no workgroup barrier per column; then the
// remaining rows are updated from contiguous row segments of A.  The sweeps
// keep the reference's orders inside a block: back_substitution /
// forward_substitution divide by the stored diagonal,
// transpose_back_substitution divides and then subtracts x[i] * l[i][j]
// (cholesky.rs:329-361).
template <class T, int MODE>
__device__ void tri_solve(const T *__restrict__ A, int n, T *__restrict__ X, int nrhs, Smem &sm) {
    constexpr int TB = 32, DL = TB + 1;
    T *D = reinterpret_cast<T *>(sm.panel);
    T *XB = D + TB * DL;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool fwd = MODE == TS_LOWER_UNIT || MODE == TS_LOWER;
    const int nblk = (n + TB - 1) / TB;
    for (int bi = 0; bi < nblk; ++bi) {
        const int blk = fwd ? bi : nblk - 1 - bi;
        const int b0 = blk * TB, bs = min(TB, n - b0);
        {   // stage the diagonal block and the block of X (all loads first)
            constexpr int PER = TB * TB / BTH;
            T v[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = tid + q * BTH, r = e / TB, c = e % TB;
                v[q] = (r < bs && c < bs) ? A[(long)(b0 + r) * n + b0 + c] : T(0);
            }
            T xv[2];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int e = tid + q * BTH, r = e / TB, i = e % TB;
                xv[q] = (r < nrhs && i < bs) ? X[(long)r * n + b0 + i] : T(0);
            }
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int e = tid + q * BTH;
                D[(e / TB) * DL + e % TB] = v[q];
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int e = tid + q * BTH;
                if (e / TB < nrhs) XB[e] = xv[q];
            }
        }
        __syncthreads();
        {
#pragma clang fp contract(off)
            const int i = lane & 31, half = lane >> 5;
            for (int r0 = 2 * wave; r0 < nrhs; r0 += 2 * BNW) {
                const int r = r0 + half;
                T x = r < nrhs ? XB[r * TB + i] : T(0);
                if (fwd) {
                    for (int c = 0; c < bs; ++c) {
                        const T l = D[i * DL + c];
                        if (MODE == TS_LOWER && i == c) x = x / D[c * DL + c];
                        const T lo = lane_get(x, c), hi = lane_get(x, 32 + c);
                        const T xc = half ? hi : lo;
                        if (i > c) x = x - l * xc;
                    }
                } else {
                    for (int c = bs - 1; c >= 0; --c) {
                        const T l = MODE == TS_UPPER ? D[i * DL + c] : D[c * DL + i];
                        if (i == c) x = x / D[c * DL + c];
                        const T lo = lane_get(x, c), hi = lane_get(x, 32 + c);
                        const T xc = half ? hi : lo;
                        if (i < c) x = x - (MODE == TS_UPPER ? l * xc : xc * l);
                    }
                }
                if (r < nrhs && i < bs) XB[r * TB + i] = x;
            }
        }
        __syncthreads();
        for (int e = tid; e < nrhs * bs; e += BTH) {
            const int r = e / bs, i = e - r * bs;
            X[(long)r * n + b0 + i] = XB[r * TB + i];
        }
        // the rest of the vector (the block's entries of A loaded first)
        const int lo = fwd ? b0 + bs : 0, hi = fwd ? n : b0;
        const int cnt = hi - lo;
        for (int q = tid; q < cnt * nrhs; q += BTH) {
            const int r = q / cnt, i = lo + (q - r * cnt);
            T av[TB];
#pragma unroll
            for (int c = 0; c < TB; ++c)
                if (c < bs) av[c] = MODE == TS_LOWER_T ? A[(long)(b0 + c) * n + i] : A[(long)i * n + b0 + c];
            T s = X[(long)r * n + i];
#pragma unroll
            for (int c = 0; c < TB; ++c)
                if (c < bs) s = s - av[c] * XB[r * TB + c];
            X[(long)r * n + i] = s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- Cholesky

// Left-looking Cholesky of A in place (lower triangle), RHS forward-solved
// along the way (B := L^-1 B).  Returns the first failing column or n.
template <class T, int NB>
__device__ int chol_fused(T *__restrict__ A, int n, T *__restrict__ B, int nrhs, Smem &sm) {
    T *P = reinterpret_cast<T *>(sm.panel);
    constexpr int LD = NB + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) sm.info = n;
    __syncthreads();
    for (int k = 0; k < n; k += NB) {
        const int kb = min(NB, n - k), m = n - k;
        panel_move<T, LD, false, true>(A, n, k, m, kb, P);
        __syncthreads();
        {
#pragma clang fp contract(off)
            // diagonal block, column by column (cholesky.rs:135-163 within the block)
            for (int j = 0; j < kb; ++j) {
                for (int r = j + tid; r < m; r += BTH) {
                    if (j > 0) {
                        // utils::dot order over the block's columns 0..j
                        T p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
                        int off = 0;
                        for (; j - off >= 8; off += 8) {
                            p0 = p0 + P[r * LD + off + 0] * P[j * LD + off + 0];
                            p1 = p1 + P[r * LD + off + 1] * P[j * LD + off + 1];
                            p2 = p2 + P[r * LD + off + 2] * P[j * LD + off + 2];
                            p3 = p3 + P[r * LD + off + 3] * P[j * LD + off + 3];
                            p4 = p4 + P[r * LD + off + 4] * P[j * LD + off + 4];
                            p5 = p5 + P[r * LD + off + 5] * P[j * LD + off + 5];
                            p6 = p6 + P[r * LD + off + 6] * P[j * LD + off + 6];
                            p7 = p7 + P[r * LD + off + 7] * P[j * LD + off + 7];
                        }
                        T s = T(0);
                        s = s + p0 + p4;
                        s = s + p1 + p5;
                        s = s + p2 + p6;
                        s = s + p3 + p7;
                        for (; off < j; ++off) s = s + P[r * LD + off] * P[j * LD + off];
                        P[r * LD + j] = P[r * LD + j] - s;
                    }
                }
                // the diagonal entry (row j is thread 0's) kept apart: the
                // division below rewrites it while other threads read it
                if (tid == 0) sm.div[j] = (double)P[j * LD + j];
                __syncthreads();
                // every thread takes the square root of the diagonal itself
                const T d = (T)sm.div[j];
                const T dv = sqrt(d);
                if (tid == 0 && (fabs(d) < beps<T>() || d < T(0)) && sm.info == n) sm.info = k + j;
                for (int r = j + tid; r < m; r += BTH) P[r * LD + j] = P[r * LD + j] / dv;
                __syncthreads();
            }
        }
        panel_move<T, LD, true, true>(A, n, k, m, kb, P);
        // RHS: y1 = L11^-1 b1 (forward_substitution order per vector), then
        // b2 -= L21 y1
        // (the block of each vector in registers: solved in place in memory,
        // every read would wait for the previous step's store)
        for (int r = tid; r < nrhs; r += BTH) {
#pragma clang fp contract(off)
            T *x = B + (long)r * n + k;
            T xv[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) xv[i] = i < kb ? x[i] : T(0);
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                if (i < kb) {
                    T s = T(0);
#pragma unroll
                    for (int c = 0; c < i; ++c) s = s + P[i * LD + c] * xv[c];
                    xv[i] = (xv[i] - s) / P[i * LD + i];
                }
            }
#pragma unroll
            for (int i = 0; i < NB; ++i)
                if (i < kb) x[i] = xv[i];
        }
        __syncthreads();
        const int mt = m - kb;
        if (mt > 0) {
            for (int q = tid; q < mt * nrhs; q += BTH) {
                const int r = q / mt, i = q - r * mt;
                T *br = B + (long)r * n;
                T yv[NB];
#pragma unroll
                for (int c = 0; c < NB; ++c) yv[c] = c < kb ? br[k + c] : T(0);
                T s = br[k + kb + i];
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < kb) s = s - P[(kb + i) * LD + c] * yv[c];
                br[k + kb + i] = s;
            }
            // SYRK on lower tiles: A22 -= L21 L21^T, in strips of TSTRIP tiles
            // along a tile row (tiles right of the diagonal tile skipped), the
            // strip's C tiles loaded before its MFMAs
            using Acc = typename BMfma<T>::acc;
            constexpr int KS = NB / 4;
            const int tiles = (mt + 15) / 16;
            const int lr = lane & 15, lk = lane >> 4;
            T *C = A + (long)(k + kb) * n + k + kb;
            int t = 0;
            for (int ti = 0; ti < tiles; ++ti) {
                T af[KS];
#pragma unroll
                for (int q = 0; q < KS; ++q) {
                    const int kk = 4 * q + lk, ai = ti * 16 + lr;
                    af[q] = (ai < mt && kk < kb) ? -P[(kb + ai) * LD + kk] : T(0);
                }
                for (int tj0 = 0; tj0 <= ti; tj0 += TSTRIP, ++t) {
                    if (t % BNW != wave) continue;
                    Acc acc[TSTRIP];
#pragma unroll
                    for (int s = 0; s < TSTRIP; ++s) {
                        const int gj = (tj0 + s) * 16 + lr;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int gi = ti * 16 + BMfma<T>::row(lk, r);
                            acc[s][r] = (tj0 + s <= ti && gi < mt && gj < mt) ? C[(long)gi * n + gj] : T(0);
                        }
                    }
#pragma unroll
                    for (int s = 0; s < TSTRIP; ++s) {
                        const int gj = (tj0 + s) * 16 + lr;
#pragma unroll
                        for (int q = 0; q < KS; ++q) {
                            const int kk = 4 * q + lk;
                            const T bf = (gj < mt && kk < kb) ? P[(kb + gj) * LD + kk] : T(0);
                            acc[s] = BMfma<T>::op(af[q], bf, acc[s]);
                        }
                    }
#pragma unroll
                    for (int s = 0; s < TSTRIP; ++s) {
                        const int gj = (tj0 + s) * 16 + lr;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int gi = ti * 16 + BMfma<T>::row(lk, r);
                            if (tj0 + s <= ti && gi < mt && gj < mt && gj <= gi) C[(long)gi * n + gj] = acc[s][r];
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    return sm.info;
}

// ---------------------------------------------------------------- dispatch

template <class T>
__device__ int solve_system(int kind, int nb, T *A, int n, T *B, int nrhs, int64_t *ipiv, Smem &sm) {
    int info;
    if (kind == RLG_KIND_LU) {
        if (nb == 8) info = lu_fused<T, 8>(A, n, B, nrhs, ipiv, sm);
        else if (nb == 32) info = lu_fused<T, 32>(A, n, B, nrhs, ipiv, sm);
        else info = lu_fused<T, 16>(A, n, B, nrhs, ipiv, sm);
        if (info == n) tri_solve<T, TS_UPPER>(A, n, B, nrhs, sm);
    } else {
        if (nb == 8) info = chol_fused<T, 8>(A, n, B, nrhs, sm);
        else if (nb == 32) info = chol_fused<T, 32>(A, n, B, nrhs, sm);
        else info = chol_fused<T, 16>(A, n, B, nrhs, sm);
        if (info == n) tri_solve<T, TS_LOWER_T>(A, n, B, nrhs, sm);
    }
    return info;
}

__global__ void k_fill_u32_graph(unsigned *p, unsigned count) {
    if (threadIdx.x < count) p[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(BTH) void k_batch_f64(rlg_batch_item *__restrict__ items,
                                                   const int *__restrict__ order,
                                                   const unsigned char *__restrict__ nbs, int count,
                                                   unsigned *__restrict__ counter) {
    __shared__ Smem sm;
    for (;;) {
        if (threadIdx.x == 0) sm.item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int idx = __builtin_amdgcn_readfirstlane(sm.item);
        __syncthreads();
        if (idx >= count) break;
        const int id = order[idx];
        rlg_batch_item it = items[id];
        const int info = solve_system<double>(it.kind, nbs[id], it.a, it.n, it.b, it.nrhs, it.ipiv, sm);
        if (threadIdx.x == 0)
            items[id].status = info == it.n ? RLG_OK
                               : (it.kind == RLG_KIND_LU ? RLG_E_DIV_BY_ZERO : RLG_E_DECOMP_FAILURE);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- mixed

// Per-workgroup workspace for the mixed form: FP32 factor (n*n), FP32
// correction / pivots, FP64 residual and iterate, FP64 fallback copy.
struct MixedSlot {
    float *a32;      // BMAXN * BMAXN
    float *d32;      // BMAXR * BMAXN
    double *res;     // BMAXR * BMAXN
    double *x;       // BMAXR * BMAXN
    double *a64;     // BMAXN * BMAXN (fallback)
    int64_t *piv;    // BMAXN
};
constexpr size_t SLOT_BYTES = (size_t)BMAXN * BMAXN * 4 + (size_t)BMAXR * BMAXN * 4 +
                              2 * (size_t)BMAXR * BMAXN * 8 + (size_t)BMAXN * BMAXN * 8 + BMAXN * 8;

__device__ MixedSlot slot_of(char *ws, int s) {
    char *p = ws + (size_t)s * SLOT_BYTES;
    MixedSlot m;
    m.a32 = reinterpret_cast<float *>(p); p += (size_t)BMAXN * BMAXN * 4;
    m.d32 = reinterpret_cast<float *>(p); p += (size_t)BMAXR * BMAXN * 4;
    m.res = reinterpret_cast<double *>(p); p += (size_t)BMAXR * BMAXN * 8;
    m.x = reinterpret_cast<double *>(p); p += (size_t)BMAXR * BMAXN * 8;
    m.a64 = reinterpret_cast<double *>(p); p += (size_t)BMAXN * BMAXN * 8;
    m.piv = reinterpret_cast<int64_t *>(p);
    return m;
}

// Sum and maximum over the wave (uniform results, fixed order).
__device__ __forceinline__ double wave_sum_f64(double v) {
    v = row16_sum(v);
    return (lane_get(v, 0) + lane_get(v, 16)) + (lane_get(v, 32) + lane_get(v, 48));
}
__device__ __forceinline__ double wave_max_f64(double v) {
    v = fmax(v, dpp<DPP_XOR1>(v));
    v = fmax(v, dpp<DPP_XOR2>(v));
    v = fmax(v, dpp<DPP_HALF_MIRROR>(v));
    v = fmax(v, dpp<DPP_MIRROR>(v));
    return fmax(fmax(lane_get(v, 0), lane_get(v, 16)), fmax(lane_get(v, 32), lane_get(v, 48)));
}

__global__ __launch_bounds__(BTH) void k_batch_mixed(rlg_batch_item *__restrict__ items,
                                                     const int *__restrict__ order,
                                                     const unsigned char *__restrict__ nbs, int count,
                                                     unsigned *__restrict__ counter, char *ws,
                                                     int max_iter, double tol,
                                                     unsigned *__restrict__ fallbacks) {
    __shared__ Smem sm;
    const MixedSlot ms = slot_of(ws, blockIdx.x);
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) sm.item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int idx = __builtin_amdgcn_readfirstlane(sm.item);
        __syncthreads();
        if (idx >= count) break;
        const int id = order[idx];
        rlg_batch_item it = items[id];
        const int n = it.n, nrhs = it.nrhs;
        const double *A = it.a;
        const int lane = tid & 63, wave = tid >> 6;
        // per-vector norms and convergence state in the panel buffer past
        // what the triangular solves (first 1568 doubles) and the row map use
        double *bmax = sm.panel + 6144, *xmax = bmax + BMAXR, *rmax = xmax + BMAXR;  // rmax[BNW][BMAXR]
        double *etab = rmax + BNW * BMAXR;                                            // etab[2][BMAXR]
        // FP32 copy (eight loads in flight), ||A||inf one row per wave
        for (int e0 = tid; e0 < n * n; e0 += 8 * BTH) {
            double v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = e0 + q * BTH < n * n ? A[e0 + q * BTH] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (e0 + q * BTH < n * n) ms.a32[e0 + q * BTH] = (float)v[q];
        }
        {   // four rows per wave step, so four rows' loads are in flight at once
            double rs = 0.0;
            for (int i0 = 4 * wave; i0 < n; i0 += 4 * BNW) {
                double s[4] = {0.0, 0.0, 0.0, 0.0};
                for (int j = lane; j < n; j += 64) {
                    double a[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) a[q] = i0 + q < n ? A[(long)(i0 + q) * n + j] : 0.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) s[q] += fabs(a[q]);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (i0 + q < n) rs = fmax(rs, wave_sum_f64(s[q]));
            }
            if (lane == 0) sm.red_v[wave] = rs;
        }
        __syncthreads();
        double anorm = sm.red_v[0];
#pragma unroll
        for (int q = 1; q < BNW; ++q) anorm = fmax(anorm, sm.red_v[q]);
        __syncthreads();
        int info;
        if (it.kind == RLG_KIND_LU) {
            const int nb = nbs[id];
            if (nb == 8) info = lu_fused<float, 8>(ms.a32, n, nullptr, 0, ms.piv, sm);
            else if (nb == 32) info = lu_fused<float, 32>(ms.a32, n, nullptr, 0, ms.piv, sm);
            else info = lu_fused<float, 16>(ms.a32, n, nullptr, 0, ms.piv, sm);
        } else {
            const int nb = nbs[id];
            if (nb == 8) info = chol_fused<float, 8>(ms.a32, n, nullptr, 0, sm);
            else if (nb == 32) info = chol_fused<float, 32>(ms.a32, n, nullptr, 0, sm);
            else info = chol_fused<float, 16>(ms.a32, n, nullptr, 0, sm);
        }
        // row permutation of the FP32 factor: raw[k] = original row at k
        int *raw = reinterpret_cast<int *>(sm.panel + 4096);
        if (tid == 0 && it.kind == RLG_KIND_LU && info == n) {
            for (int i = 0; i < n; ++i) raw[i] = i;
            for (int k = 0; k < n; ++k) {
                const int p = (int)ms.piv[k];
                if (p != k) { const int t = raw[k]; raw[k] = raw[p]; raw[p] = t; }
            }
        }
        // refinement of all right-hand sides together: x = 0, r = b, |b|
        double *X = ms.x, *R = ms.res;
        float *D = ms.d32;
        for (int r = wave; r < nrhs; r += BNW) {
            double bm = 0.0;
            for (int i = lane; i < n; i += 64) {
                const double bv = it.b[(long)r * n + i];
                X[(long)r * n + i] = 0.0;
                R[(long)r * n + i] = bv;
                bm = fmax(bm, fabs(bv));
            }
            bm = wave_max_f64(bm);
            if (lane == 0) bmax[r] = bm;
        }
        __syncthreads();
        // done: vectors whose backward error reached tol (the same bits in
        // every thread); a vector that stops halving its error fails the system
        unsigned done = 0u;
        const unsigned all = nrhs >= 32 ? ~0u : (1u << nrhs) - 1u;
        bool stalled = info != n;
        for (int itn = 0; itn < max_iter && !stalled && done != all; ++itn) {
            for (int e = tid; e < nrhs * n; e += BTH) {
                const int r = e / n, i = e - r * n;
                const int src = it.kind == RLG_KIND_LU ? raw[i] : i;
                D[e] = (done >> r) & 1u ? 0.f : (float)R[(long)r * n + src];
            }
            __syncthreads();
            if (it.kind == RLG_KIND_LU) {
                tri_solve<float, TS_LOWER_UNIT>(ms.a32, n, D, nrhs, sm);
                tri_solve<float, TS_UPPER>(ms.a32, n, D, nrhs, sm);
            } else {
                tri_solve<float, TS_LOWER>(ms.a32, n, D, nrhs, sm);
                tri_solve<float, TS_LOWER_T>(ms.a32, n, D, nrhs, sm);
            }
            // x += d and |x|, one vector per wave
            for (int r = wave; r < nrhs; r += BNW) {
                if ((done >> r) & 1u) continue;
                double xm = 0.0;
                for (int i = lane; i < n; i += 64) {
                    const double v = X[(long)r * n + i] + (double)D[(long)r * n + i];
                    X[(long)r * n + i] = v;
                    xm = fmax(xm, fabs(v));
                }
                xm = wave_max_f64(xm);
                if (lane == 0) xmax[r] = xm;
            }
            __syncthreads();
            // r = b - A x in FP64: a wave takes four rows and four vectors per
            // pass over the rows (lanes across the row), so four rows' loads
            // are in flight at once; |r| per vector kept in the wave's own LDS
            // slots.  Each (row, vector) sum keeps one fixed order.
            if (lane == 0) {
#pragma unroll
                for (int r = 0; r < BMAXR; ++r) rmax[wave * BMAXR + r] = 0.0;
            }
            for (int i0 = 4 * wave; i0 < n; i0 += 4 * BNW) {
                for (int rb = 0; rb < nrhs; rb += 4) {
                    double acc[4][4];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[u][q] = 0.0;
                    for (int j = lane; j < n; j += 64) {
                        double a[4], xv[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) a[u] = i0 + u < n ? A[(long)(i0 + u) * n + j] : 0.0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) xv[q] = rb + q < nrhs ? X[(long)(rb + q) * n + j] : 0.0;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int q = 0; q < 4; ++q) acc[u][q] += a[u] * xv[q];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            if (i0 + u < n && rb + q < nrhs) {
                                const int i = i0 + u, r = rb + q;
                                const double res = it.b[(long)r * n + i] - wave_sum_f64(acc[u][q]);
                                if (lane == 0) {
                                    R[(long)r * n + i] = res;
                                    rmax[wave * BMAXR + r] = fmax(rmax[wave * BMAXR + r], fabs(res));
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // every thread takes the same decisions; the errors of this step
            // go to etab[itn & 1], the previous step's are read from the other
            for (int r = 0; r < nrhs; ++r) {
                if ((done >> r) & 1u) continue;
                double rn = rmax[r];
#pragma unroll
                for (int q = 1; q < BNW; ++q) rn = fmax(rn, rmax[q * BMAXR + r]);
                const double den = anorm * xmax[r] + bmax[r];
                const double eta = den > 0 ? rn / den : 0.0;
                if (eta <= tol) {
                    done |= 1u << r;
                } else if (itn >= 1 && !(eta <= 0.5 * etab[((itn + 1) & 1) * BMAXR + r])) {
                    stalled = true;
                }
                if (tid == 0) etab[(itn & 1) * BMAXR + r] = eta;
            }
            __syncthreads();
        }
        const bool converged = !stalled && done == all;
        if (converged) {
            for (int e = tid; e < nrhs * n; e += BTH) it.b[e] = ms.x[e];
            if (tid == 0) items[id].status = RLG_OK;
        } else {
            // fallback: plain FP64 on a copy of A
            for (int e = tid; e < n * n; e += BTH) ms.a64[e] = A[e];
            __syncthreads();
            const int inf64 = solve_system<double>(it.kind, nbs[id], ms.a64, n, it.b, nrhs, ms.piv, sm);
            if (tid == 0) {
                items[id].status = inf64 == n ? RLG_OK
                                   : (it.kind == RLG_KIND_LU ? RLG_E_DIV_BY_ZERO : RLG_E_DECOMP_FAILURE);
                atomicAdd(fallbacks, 1u);
            }
        }
        __syncthreads();
    }
}

}  // namespace rlg
