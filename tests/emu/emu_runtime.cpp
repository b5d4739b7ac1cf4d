// Fiber scheduler and wave collectives of the CPU emulation (test
// infrastructure).  Every lane of the workgroup is a fiber on the calling
// thread; the rules are stated in hip_emu.hpp.
#include "emu_runtime.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#endif
#endif
#ifdef EMU_ASAN
#include <sanitizer/common_interface_defs.h>
#else
static inline void __sanitizer_start_switch_fiber(void **, const void *, size_t) {}
static inline void __sanitizer_finish_switch_fiber(void *, const void **, size_t *) {}
#endif

// Context switch between fibers (x86-64 System V): push the callee-saved
// registers, store the stack pointer in *save, continue on the stack `next`
// (which another call left in the same state, or which start_stack built).
extern "C" void emu_switch(void **save, void *next);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

namespace emu {

Dim3 g_thread, g_block, g_block_dim, g_grid_dim;

namespace {

enum State : unsigned char { RUN, BAR, COL, DONE };
enum Kind : unsigned char { K_READLANE, K_FIRST, K_DPP, K_BALLOT, K_MFMA64, K_MFMA32 };
const char *kind_name(int k) {
    static const char *names[] = {"readlane", "readfirstlane", "dpp", "ballot", "mfma_f64", "mfma_f32"};
    return names[k];
}

constexpr size_t STACK = 64 << 10;

struct Lane {
    void *sp = nullptr;  // saved stack pointer while suspended
    char *stack = nullptr;
    void *fake = nullptr;
    State st = RUN;
    Kind kind = K_READLANE;
    int block = 0, tid = 0;
    void *site = nullptr;
    int iv = 0, old = 0, sel = 0, ctrl = 0, masks = 0;
    bool pred = false;
    double a = 0, b = 0;
    d4 c = {0, 0, 0, 0};
    float af = 0, bf = 0;
    f4 cf = {0, 0, 0, 0};
    long long ires = 0;
    d4 d = {0, 0, 0, 0};
    f4 df = {0, 0, 0, 0};
};

struct Run {
    std::vector<Lane> lanes;
    void *main_sp = nullptr;
    void *main_fake = nullptr;
    const void *main_bottom = nullptr;
    size_t main_size = 0;
    int cur = -1;
    int threads = 0, blocks = 0, block_id = 0;
    void (*body)(void *) = nullptr;
    void *arg = nullptr;
    std::string err;
    Stats stats;
};

Run *g_run = nullptr;

Lane &cur() { return g_run->lanes[g_run->cur]; }

void fail(const std::string &msg) {
    if (g_run->err.empty()) g_run->err = msg;
}

// Lane -> scheduler.  Returns when the scheduler resumes the lane.
void yield(State st) {
    Lane &L = cur();
    L.st = st;
    __sanitizer_start_switch_fiber(&L.fake, g_run->main_bottom, g_run->main_size);
    emu_switch(&L.sp, g_run->main_sp);
    __sanitizer_finish_switch_fiber(L.fake, nullptr, nullptr);
}

// First code a fiber runs; it never returns (it ends by switching away).
void fiber_entry() {
    __sanitizer_finish_switch_fiber(nullptr, &g_run->main_bottom, &g_run->main_size);
    g_run->body(g_run->arg);
    Lane &L = cur();
    L.st = DONE;
    __sanitizer_start_switch_fiber(nullptr, g_run->main_bottom, g_run->main_size);
    emu_switch(&L.sp, g_run->main_sp);
    std::abort();  // a finished fiber is never resumed
}

// A fresh stack laid out as emu_switch leaves a suspended one: six register
// slots, then fiber_entry as the return address, then a dummy return
// address so that fiber_entry starts with the stack alignment of a call.
void *start_stack(char *stack, size_t size) {
    auto top = reinterpret_cast<uintptr_t>(stack + size) & ~uintptr_t(15);
    void **sp = reinterpret_cast<void **>(top);
    *--sp = nullptr;
    *--sp = reinterpret_cast<void *>(&fiber_entry);
    for (int i = 0; i < 6; ++i) *--sp = nullptr;
    return sp;
}

void resume(Run &R, int idx) {
    Lane &L = R.lanes[idx];
    R.cur = idx;
    g_thread.x = (unsigned)L.tid;
    g_block.x = (unsigned)R.block_id;
    ++R.stats.switches;
    __sanitizer_start_switch_fiber(&R.main_fake, L.stack, STACK);
    emu_switch(&R.main_sp, L.sp);
    __sanitizer_finish_switch_fiber(R.main_fake, nullptr, nullptr);
}

std::string where(const Lane &L) {
    char buf[96];
    snprintf(buf, sizeof buf, "block %d thread %d (wave %d lane %d)", g_run->block_id, L.tid, L.tid / 64,
             L.tid % 64);
    return buf;
}

int dpp_source(int ctrl, int l, bool *valid) {
    *valid = true;
    if (ctrl >= 0x00 && ctrl <= 0xFF) return (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);  // quad_perm
    if (ctrl >= 0x121 && ctrl <= 0x12F) return (l & ~15) | ((l - (ctrl - 0x120)) & 15);  // row_ror
    if (ctrl == 0x140) return (l & ~15) | (15 - (l & 15));                                // row_mirror
    if (ctrl == 0x141) return (l & ~7) | (7 - (l & 7));                                   // row_half_mirror
    if (ctrl >= 0x111 && ctrl <= 0x11F) {                                                 // row_shr
        const int s = (l & 15) - (ctrl - 0x110);
        *valid = s >= 0;
        return (l & ~15) | (s & 15);
    }
    if (ctrl >= 0x101 && ctrl <= 0x10F) {                                                 // row_shl
        const int s = (l & 15) + (ctrl - 0x100);
        *valid = s < 16;
        return (l & ~15) | (s & 15);
    }
    *valid = false;
    return -1;
}

// All 64 lanes of the wave starting at `base` wait at one collective.
void compute_collective(Run &R, int base) {
    Lane *w = &R.lanes[base];
    const Kind k = w[0].kind;
    switch (k) {
    case K_READLANE: {
        const int s = w[0].sel;
        if (s < 0 || s > 63) { fail("readlane of lane " + std::to_string(s) + " at " + where(w[0])); return; }
        for (int l = 0; l < 64; ++l) w[l].ires = w[s].iv;
        break;
    }
    case K_FIRST:
        for (int l = 0; l < 64; ++l) w[l].ires = w[0].iv;
        break;
    case K_DPP: {
        if (w[0].masks != 0xFF) { fail("DPP with row/bank masks other than 0xF at " + where(w[0])); return; }
        int res[64];
        for (int l = 0; l < 64; ++l) {
            bool ok;
            const int s = dpp_source(w[0].ctrl, l, &ok);
            if (s < 0) { fail("unsupported DPP control " + std::to_string(w[0].ctrl) + " at " + where(w[0])); return; }
            res[l] = ok ? w[s].iv : w[l].old;
        }
        for (int l = 0; l < 64; ++l) w[l].ires = res[l];
        break;
    }
    case K_BALLOT: {
        unsigned long long m = 0;
        for (int l = 0; l < 64; ++l)
            if (w[l].pred) m |= 1ull << l;
        for (int l = 0; l < 64; ++l) w[l].ires = (long long)m;
        break;
    }
    case K_MFMA64: {
        // lane l = 16 lk + lr holds A[lr][lk] and B[lk][lr]; accumulator
        // register r of lane l is D[lk + 4 r][lr]
        for (int l = 0; l < 64; ++l) {
            const int lr = l & 15, lk = l >> 4;
            d4 d;
            for (int r = 0; r < 4; ++r) {
                const int i = lk + 4 * r;
                double acc = w[l].c[r];
                for (int kk = 0; kk < 4; ++kk) acc = std::fma(w[kk * 16 + i].a, w[kk * 16 + lr].b, acc);
                d[r] = acc;
            }
            w[l].d = d;
        }
        break;
    }
    case K_MFMA32: {
        // as FP64, with accumulator register r of lane l holding D[4 lk + r][lr]
        for (int l = 0; l < 64; ++l) {
            const int lr = l & 15, lk = l >> 4;
            f4 d;
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * lk + r;
                float acc = w[l].cf[r];
                for (int kk = 0; kk < 4; ++kk) acc = std::fma(w[kk * 16 + i].af, w[kk * 16 + lr].bf, acc);
                d[r] = acc;
            }
            w[l].df = d;
        }
        break;
    }
    }
    ++R.stats.collectives;
    for (int l = 0; l < 64; ++l) w[l].st = RUN;
}

// Releases what can be released; returns whether any lane became runnable.
bool resolve(Run &R) {
    bool progressed = false;
    const int T = R.threads;
    for (int b = 0; b < R.blocks; ++b) {
        for (int w0 = 0; w0 < T; w0 += 64) {
            const int base = b * T + w0;
            int col = 0, bar = 0, done = 0;
            for (int l = 0; l < 64; ++l) {
                const State s = R.lanes[base + l].st;
                col += s == COL;
                bar += s == BAR;
                done += s == DONE;
            }
            if (col == 0) continue;
            if (col != 64) {
                fail(std::string("a ") + kind_name(R.lanes[base].kind) + " reached by " + std::to_string(col) +
                     " of 64 lanes (" + std::to_string(bar) + " at a barrier, " + std::to_string(done) +
                     " exited) in " + where(R.lanes[base]));
                return false;
            }
            const Lane &f = R.lanes[base];
            for (int l = 1; l < 64; ++l) {
                const Lane &x = R.lanes[base + l];
                if (x.site != f.site || x.kind != f.kind) {
                    fail(std::string("lanes of one wave at different collectives: ") + kind_name(f.kind) +
                         " / " + kind_name(x.kind) + " in " + where(x));
                    return false;
                }
                if ((f.kind == K_READLANE && x.sel != f.sel) ||
                    (f.kind == K_DPP && (x.ctrl != f.ctrl || x.masks != f.masks))) {
                    fail(std::string("non-uniform operand of ") + kind_name(f.kind) + " in " + where(x));
                    return false;
                }
            }
            compute_collective(R, base);
            if (!R.err.empty()) return false;
            progressed = true;
        }
        int bar = 0, live = 0;
        void *site = nullptr;
        for (int t = 0; t < T; ++t) {
            const Lane &x = R.lanes[b * T + t];
            if (x.st == DONE) continue;
            ++live;
            if (x.st == BAR) {
                if (bar && x.site != site) {
                    fail("threads of one workgroup at different barriers: " + where(x));
                    return false;
                }
                site = x.site;
                ++bar;
            }
        }
        if (live && bar == live) {
            for (int t = 0; t < T; ++t)
                if (R.lanes[b * T + t].st == BAR) R.lanes[b * T + t].st = RUN;
            ++R.stats.barriers;
            progressed = true;
        }
    }
    return progressed;
}

}  // namespace

// ------------------------------------------------------------- lane side

void syncthreads() {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    yield(BAR);
}

int readlane(int v, int lane) {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    L.kind = K_READLANE;
    L.iv = v;
    L.sel = lane;
    yield(COL);
    return (int)L.ires;
}

int readfirstlane(int v) {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    L.kind = K_FIRST;
    L.iv = v;
    yield(COL);
    return (int)L.ires;
}

int update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    L.kind = K_DPP;
    L.iv = src;
    L.old = bound_ctrl ? 0 : old;
    L.ctrl = ctrl;
    L.masks = ((row_mask & 0xF) << 4) | (bank_mask & 0xF);
    yield(COL);
    return (int)L.ires;
}

unsigned long long ballot(bool pred) {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    L.kind = K_BALLOT;
    L.pred = pred;
    yield(COL);
    return (unsigned long long)L.ires;
}

d4 mfma_f64_16x16x4(double a, double b, d4 c) {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    L.kind = K_MFMA64;
    L.a = a;
    L.b = b;
    L.c = c;
    yield(COL);
    return L.d;
}

f4 mfma_f32_16x16x4(float a, float b, f4 c) {
    Lane &L = cur();
    L.site = __builtin_return_address(0);
    L.kind = K_MFMA32;
    L.af = a;
    L.bf = b;
    L.cf = c;
    yield(COL);
    return L.df;
}

// ------------------------------------------------------------- launch

int launch(int block_id, int grid, int threads, int order, unsigned seed, void (*body)(void *), void *arg,
           std::string *err, Stats *stats) {
    if (grid < 1 || block_id < 0 || block_id >= grid || threads < 64 || threads % 64) {
        *err = "bad launch shape";
        return 1;
    }
    Run R;
    R.blocks = 1;
    R.threads = threads;
    R.body = body;
    R.arg = arg;
    R.block_id = block_id;
    R.lanes.resize((size_t)threads);
    g_run = &R;
    g_block_dim = Dim3{(unsigned)threads, 1, 1};
    g_grid_dim = Dim3{(unsigned)grid, 1, 1};
    for (int i = 0; i < (int)R.lanes.size(); ++i) {
        Lane &L = R.lanes[i];
        L.block = i / threads;
        L.tid = i % threads;
        L.stack = static_cast<char *>(std::malloc(STACK));
        L.sp = start_stack(L.stack, STACK);
    }
    std::vector<int> seq(R.lanes.size());
    for (int i = 0; i < (int)seq.size(); ++i) seq[i] = i;
    if (order == 1) std::reverse(seq.begin(), seq.end());
    std::mt19937 rng(seed);
    for (;;) {
        if (order == 2) std::shuffle(seq.begin(), seq.end(), rng);
        for (int idx : seq) {
            if (R.lanes[idx].st == RUN) resume(R, idx);
            if (!R.err.empty()) break;
        }
        if (!R.err.empty()) break;
        bool all_done = true;
        for (const Lane &L : R.lanes) all_done = all_done && L.st == DONE;
        if (all_done) break;
        if (!resolve(R)) {
            if (R.err.empty()) {
                int bar = 0, col = 0;
                for (const Lane &L : R.lanes) { bar += L.st == BAR; col += L.st == COL; }
                R.err = "no lane can proceed: " + std::to_string(bar) + " at barriers, " + std::to_string(col) +
                        " at collectives";
            }
            break;
        }
    }
    // a failed run leaves suspended fibers: their stacks are simply dropped
    for (Lane &L : R.lanes) std::free(L.stack);
    if (stats) *stats = R.stats;
    g_run = nullptr;
    *err = R.err;
    return R.err.empty() ? 0 : 1;
}

}  // namespace emu
