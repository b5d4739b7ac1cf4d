// Host side of the batched factor-and-solve: item validation and ordering,
// panel-width table and autotuner, grid sizing, the eager launches, the graph
// form and the C entry points.  The kernels are in batched_device.hpp.

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "batched_device.hpp"
#include "kernels.hpp"

namespace rlg {

// ---------------------------------------------------------------- host side

namespace {

// Autotuned panel width per n-bucket (n <= 32, 64, 128, 256, 512).
std::mutex g_tune_mu;
int g_nb_table[5] = {8, 16, 16, 16, 16};

int bucket_of(int n) {
    if (n <= 32) return 0;
    if (n <= 64) return 1;
    if (n <= 128) return 2;
    if (n <= 256) return 3;
    return 4;
}

int nb_for(int n) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    int nb = g_nb_table[bucket_of(n)];
    if (nb == 32 && (n + 16) * 33 > PANEL) nb = 16;
    return nb;
}

// One workgroup per CU: the kernels take 256 VGPRs + 256 AGPRs per lane
// (one wave per SIMD) and 73 KB of LDS, so a second workgroup would only
// wait for the first to finish.
int resident_grid(Device *d) { return d->num_cus; }

struct Prepared {
    PoolBuf order, nbs, counter;
    int count = 0;
};

int prepare(hipStream_t s, size_t count, rlg_batch_item *items, Prepared &p) {
    std::vector<rlg_batch_item> h(count);
    RLG_HIP(hipMemcpyAsync(h.data(), items, count * sizeof(rlg_batch_item), hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    std::vector<int> order(count);
    std::vector<unsigned char> nbs(count);
    for (size_t i = 0; i < count; ++i) {
        const auto &it = h[i];
        if (it.n < 1 || it.n > BMAXN || it.nrhs < 0 || it.nrhs > BMAXR || !it.a ||
            (it.nrhs > 0 && !it.b) || (it.kind != RLG_KIND_LU && it.kind != RLG_KIND_CHOL) ||
            (it.kind == RLG_KIND_LU && !it.ipiv)) {
            set_error("batch item %zu is invalid (kind %d, n %d, nrhs %d)", i, it.kind, it.n, it.nrhs);
            return RLG_E_INVALID_ARG;
        }
        order[i] = (int)i;
        nbs[i] = (unsigned char)nb_for(it.n);
    }
    // longest first: cost ~ n^3 (Cholesky about half of LU)
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const double cx = (double)h[x].n * h[x].n * h[x].n * (h[x].kind == RLG_KIND_LU ? 2 : 1);
        const double cy = (double)h[y].n * h[y].n * h[y].n * (h[y].kind == RLG_KIND_LU ? 2 : 1);
        return cx > cy;
    });
    RLG_TRY(p.order.alloc(count * sizeof(int), s));
    RLG_TRY(p.nbs.alloc(count, s));
    RLG_TRY(p.counter.alloc(sizeof(unsigned) * 2, s));
    RLG_HIP(hipMemcpyAsync(p.order.p, order.data(), count * sizeof(int), hipMemcpyHostToDevice, s));
    RLG_HIP(hipMemcpyAsync(p.nbs.p, nbs.data(), count, hipMemcpyHostToDevice, s));
    RLG_TRY(fill_u32(p.counter.as<uint32_t>(), 2, 0u, s));
    RLG_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    p.count = (int)count;
    return RLG_OK;
}

}  // namespace

int batch_solve_f64(hipStream_t s, Device *d, size_t count, rlg_batch_item *items) {
    if (count == 0) return RLG_OK;
    Prepared p;
    RLG_TRY(prepare(s, count, items, p));
    const int grid = std::min<int>(resident_grid(d), (int)count);
    RLG_PROF("batch_f64", s);
    hipLaunchKernelGGL(k_batch_f64, dim3(grid), dim3(BTH), 0, s, items, p.order.as<int>(),
                       p.nbs.as<unsigned char>(), p.count, p.counter.as<unsigned>());
    RLG_LAUNCHED();
    return RLG_OK;
}

int batch_solve_mixed(hipStream_t s, Device *d, size_t count, rlg_batch_item *items, int max_iter,
                      double tol, size_t *fallbacks) {
    if (count == 0) { if (fallbacks) *fallbacks = 0; return RLG_OK; }
    Prepared p;
    RLG_TRY(prepare(s, count, items, p));
    const int grid = std::min<int>(resident_grid(d), (int)count);
    PoolBuf ws;
    RLG_TRY(ws.alloc(SLOT_BYTES * grid, s));
    {
        RLG_PROF("batch_mixed", s);
        hipLaunchKernelGGL(k_batch_mixed, dim3(grid), dim3(BTH), 0, s, items, p.order.as<int>(),
                           p.nbs.as<unsigned char>(), p.count, p.counter.as<unsigned>(), ws.as<char>(),
                           max_iter, tol, p.counter.as<unsigned>() + 1);
        RLG_LAUNCHED();
    }
    unsigned fb = 0;
    RLG_HIP(hipMemcpyAsync(&fb, p.counter.as<unsigned>() + 1, 4, hipMemcpyDeviceToHost, s));
    RLG_HIP(hipStreamSynchronize(s));
    if (fallbacks) *fallbacks = fb;
    return RLG_OK;
}

void batch_set_nb(int bucket, int nb) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (bucket >= 0 && bucket < 5 && (nb == 8 || nb == 16 || nb == 32)) g_nb_table[bucket] = nb;
}

// ---------------------------------------------------------------- autotuner

// Synthetic system for tuning: a[i][j] uniform in [-1, 1) from an integer
// hash, plus n on the diagonal; b uniform.  Deterministic in (seed, i, j).
__device__ __forceinline__ double hash_unit(unsigned seed, unsigned i, unsigned j) {
    unsigned h = seed * 0x9E3779B1u ^ (i * 0x85EBCA77u) ^ (j * 0xC2B2AE3Du);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return (double)h * (2.0 / 4294967296.0) - 1.0;
}

__global__ void k_gen_systems(rlg_batch_item *__restrict__ items, int count, unsigned seed) {
    const int id = blockIdx.y;
    if (id >= count) return;
    const rlg_batch_item it = items[id];
    const int n = it.n;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n + it.nrhs * n; e += gridDim.x * blockDim.x) {
        if (e < n * n) {
            const int i = e / n, j = e - i * n;
            double v = hash_unit(seed + id, i, j);
            if (i == j) v += (double)n;
            it.a[e] = v;
        } else {
            it.b[e - n * n] = hash_unit(seed + id + 7919u, e - n * n, 1u);
        }
    }
}

namespace {
std::string g_tune_report;
}

int batch_autotune(hipStream_t s, Device *d, int per_bucket) {
    static const int rep_n[5] = {24, 56, 120, 250, 500};
    if (per_bucket < 8) per_bucket = 8;
    std::string report;
    char line[160];
    for (int b = 0; b < 5; ++b) {
        const int n = rep_n[b], nr = 4, count = per_bucket;
        const size_t per = (size_t)n * n + (size_t)nr * n + n;
        PoolBuf store, itbuf;
        RLG_TRY(store.alloc(per * 8 * count, s));
        RLG_TRY(itbuf.alloc(sizeof(rlg_batch_item) * count, s));
        std::vector<rlg_batch_item> h(count);
        for (int i = 0; i < count; ++i) {
            double *base = store.as<double>() + per * i;
            h[i] = rlg_batch_item{RLG_KIND_LU, n, nr, 0, base, base + (size_t)n * n,
                                  reinterpret_cast<int64_t *>(base + (size_t)n * n + (size_t)nr * n)};
        }
        RLG_HIP(hipMemcpyAsync(itbuf.p, h.data(), sizeof(rlg_batch_item) * count, hipMemcpyHostToDevice, s));
        float best_ms = 1e30f;
        int best_nb = 16;
        std::string row;
        for (int nb : {8, 16, 32}) {
            if (nb == 32 && (n + 16) * 33 > PANEL) continue;
            batch_set_nb(b, nb);
            float ms_sum = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                hipLaunchKernelGGL(k_gen_systems, dim3(64, (unsigned)count), dim3(256), 0, s,
                                   itbuf.as<rlg_batch_item>(), count, 1234u + rep);
                RLG_LAUNCHED();
                hipEvent_t e0, e1;
                RLG_HIP(hipEventCreate(&e0));
                RLG_HIP(hipEventCreate(&e1));
                RLG_HIP(hipEventRecord(e0, s));
                int st = batch_solve_f64(s, d, (size_t)count, itbuf.as<rlg_batch_item>());
                RLG_HIP(hipEventRecord(e1, s));
                RLG_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                if (st != RLG_OK) return st;
                if (rep) ms_sum += ms;  // first repetition warms up
            }
            snprintf(line, sizeof line, " nb%d=%.3fms", nb, ms_sum);
            row += line;
            if (ms_sum < best_ms) { best_ms = ms_sum; best_nb = nb; }
        }
        batch_set_nb(b, best_nb);
        snprintf(line, sizeof line, "bucket n<=%d (n=%d, %d systems): best nb=%d;", b == 0 ? 32 : 32 << b,
                 n, count, best_nb);
        report += line + row + "\n";
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_report = report;
    return RLG_OK;
}

// ---------------------------------------------------------------- graph form

}  // namespace rlg

struct rlg_batch_graph {
    rlg::DevBuf order, nbs, counter;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t count = 0;
};

namespace rlg {

int batch_graph_create(hipStream_t s, Device *d, size_t count, rlg_batch_item *items, rlg_batch_graph *g) {
    Prepared p;
    RLG_TRY(prepare(s, count, items, p));
    // the graph keeps its own copies of the order and panel-width tables
    RLG_TRY(g->order.alloc(count * sizeof(int)));
    RLG_TRY(g->nbs.alloc(count));
    RLG_TRY(g->counter.alloc(sizeof(unsigned) * 2));
    RLG_HIP(hipMemcpyAsync(g->order.p, p.order.p, count * sizeof(int), hipMemcpyDeviceToDevice, s));
    RLG_HIP(hipMemcpyAsync(g->nbs.p, p.nbs.p, count, hipMemcpyDeviceToDevice, s));
    RLG_HIP(hipStreamSynchronize(s));
    const int grid = std::min<int>(resident_grid(d), (int)count);
    // capture: work-counter reset (a fill kernel, not a memset node) + the persistent kernel
    RLG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_fill_u32_graph, dim3(1), dim3(64), 0, s, g->counter.as<unsigned>(), 2u);
    hipLaunchKernelGGL(k_batch_f64, dim3(grid), dim3(BTH), 0, s, items, g->order.as<int>(),
                       g->nbs.as<unsigned char>(), (int)count, g->counter.as<unsigned>());
    hipError_t ce = hipStreamEndCapture(s, &g->graph);
    if (ce != hipSuccess) {
        set_error("graph capture failed: %s", hipGetErrorString(ce));
        return RLG_E_DEVICE;
    }
    RLG_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->count = count;
    return RLG_OK;
}

}  // namespace rlg

using namespace rlg;

extern "C" {

int rlg_batch_solve_f64_dev(size_t count, rlg_batch_item *items, void *stream) {
    Device *d;
    RLG_TRY(current_device(&d));
    return batch_solve_f64(pick_stream(d, stream), d, count, items);
}

int rlg_batch_solve_mixed_dev(size_t count, rlg_batch_item *items, int max_iter, double tol,
                              size_t *fallbacks, void *stream) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (max_iter < 1) max_iter = 30;
    if (!(tol > 0)) tol = 1e-13;
    return batch_solve_mixed(pick_stream(d, stream), d, count, items, max_iter, tol, fallbacks);
}

int rlg_batch_autotune(int systems_per_bucket, char *report, size_t len) {
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_TRY(batch_autotune(d->stream, d, systems_per_bucket));
    if (report && len) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        const size_t m = std::min(len - 1, g_tune_report.size());
        std::memcpy(report, g_tune_report.data(), m);
        report[m] = 0;
    }
    return RLG_OK;
}

int rlg_batch_graph_create(size_t count, rlg_batch_item *items, rlg_batch_graph **out) {
    if (!out || !items || count == 0) { set_error("bad argument"); return RLG_E_INVALID_ARG; }
    *out = nullptr;
    Device *d;
    RLG_TRY(current_device(&d));
    hipStream_t s;
    RLG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // capture on a private stream
    auto *g = new rlg_batch_graph();
    int st = batch_graph_create(s, d, count, items, g);
    (void)hipStreamDestroy(s);
    if (st != RLG_OK) { rlg_batch_graph_destroy(g); return st; }
    *out = g;
    return RLG_OK;
}

int rlg_batch_graph_launch(rlg_batch_graph *g, void *stream) {
    if (!g || !g->exec) { set_error("bad graph"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_HIP(hipGraphLaunch(g->exec, pick_stream(d, stream)));
    return RLG_OK;
}

void rlg_batch_graph_destroy(rlg_batch_graph *g) {
    if (!g) return;
    (void)hipDeviceSynchronize();
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

}  // extern "C"
