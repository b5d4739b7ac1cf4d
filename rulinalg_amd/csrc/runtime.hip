// Runtime support: thread-local errors, per-device context, the event
// profiler, small utility kernels and the device-memory C entry points.
#include <cstdarg>
#include <map>

#include "common.hpp"

namespace rlg {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *get_error() { return g_err; }

// ------------------------------------------------------------------ devices

static std::mutex g_dev_mu;
static Device *g_devs[64] = {nullptr};
static thread_local int t_device = 0;

int current_device(Device **out) {
    int id = t_device;
    {
        std::lock_guard<std::mutex> lk(g_dev_mu);
        if (!g_devs[id]) {
            int count = 0;
            hipError_t e = hipGetDeviceCount(&count);
            if (e != hipSuccess || count == 0) {
                set_error("no HIP device available (%s)", hipGetErrorString(e));
                return RLG_E_DEVICE;
            }
            if (id >= count) {
                set_error("device %d out of range (%d devices)", id, count);
                return RLG_E_INVALID_ARG;
            }
            RLG_HIP(hipSetDevice(id));
            hipDeviceProp_t prop;
            RLG_HIP(hipGetDeviceProperties(&prop, id));
            if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
                set_error("librlgpu is built for gfx950; device %d is %s", id, prop.gcnArchName);
                return RLG_E_DEVICE;
            }
            Device *d = new Device();
            d->id = id;
            d->num_cus = prop.multiProcessorCount;
            d->lds_per_block = prop.sharedMemPerBlock;
            RLG_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
            g_devs[id] = d;
        }
    }
    RLG_HIP(hipSetDevice(id));
    *out = g_devs[id];
    return RLG_OK;
}

hipStream_t pick_stream(Device *dev, void *user_stream) {
    return user_stream ? static_cast<hipStream_t>(user_stream) : dev->stream;
}

int AuxStream::create() {
    RLG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return RLG_OK;
}

AuxStream::~AuxStream() {
    if (s) (void)hipStreamDestroy(s);
}

// ------------------------------------------------------------------ profiler

namespace {
struct ProfRec {
    hipEvent_t e0, e1;
    int cls;
    double work;
};
struct ProfState {
    std::mutex mu;
    bool on = false;
    std::vector<std::string> names;
    std::vector<ProfRec> pending;
    std::vector<int64_t> launches;
    std::vector<double> ms;
    std::vector<double> work;
};
ProfState &prof() {
    static ProfState *p = new ProfState();
    return *p;
}
void prof_collect_locked(ProfState &p) {
    for (auto &r : p.pending) {
        float t = 0.f;
        if (hipEventSynchronize(r.e1) == hipSuccess && hipEventElapsedTime(&t, r.e0, r.e1) == hipSuccess) {
            p.launches[r.cls] += 1;
            p.ms[r.cls] += t;
            p.work[r.cls] += r.work;
        }
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    p.pending.clear();
}
}  // namespace

int prof_class(const char *name) {
    auto &p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    for (size_t i = 0; i < p.names.size(); ++i)
        if (p.names[i] == name) return (int)i;
    p.names.push_back(name);
    p.launches.push_back(0);
    p.ms.push_back(0.0);
    p.work.push_back(0.0);
    return (int)p.names.size() - 1;
}

bool prof_on() { return prof().on; }

ProfScope::ProfScope(int c, hipStream_t st, double w) : cls(c), s(st), work(w) {
    if (!prof().on) return;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        e0 = e1 = nullptr;
        return;
    }
    (void)hipEventRecord(e0, s);
}

ProfScope::~ProfScope() {
    if (!e0) return;
    (void)hipEventRecord(e1, s);
    auto &p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    p.pending.push_back({e0, e1, cls, work});
    if (p.pending.size() > 20000) prof_collect_locked(p);
}

// ------------------------------------------------------------------ pool

namespace {
struct PoolEntry {
    void *p;
    size_t bytes;
    int dev;
    std::vector<hipEvent_t> ready;  // every stream that used the buffer
};
std::mutex g_pool_mu;
std::vector<PoolEntry> &pool_entries() {
    static auto *v = new std::vector<PoolEntry>();
    return *v;
}
}  // namespace

int PoolBuf::alloc(size_t n, hipStream_t s) {
    release();
    if (n == 0) n = 256;
    (void)hipGetDevice(&dev);
    stream = s;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto &v = pool_entries();
        size_t best = v.size();
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i].dev == dev && v[i].bytes >= n && v[i].bytes <= 2 * n &&
                (best == v.size() || v[i].bytes < v[best].bytes))
                best = i;
        if (best != v.size()) {
            PoolEntry e = v[best];
            v.erase(v.begin() + (long)best);
            p = e.p;
            bytes = e.bytes;
            for (hipEvent_t ev : e.ready) {
                (void)hipStreamWaitEvent(s, ev, 0);
                (void)hipEventDestroy(ev);
            }
            return RLG_OK;
        }
    }
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        // give back cached buffers and retry once
        pool_trim();
        e = hipMalloc(&p, n);
    }
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RLG_E_OOM : RLG_E_DEVICE;
    }
    bytes = n;
    return RLG_OK;
}

// An event recorded on s, or (if that fails) s drained on the host.
static hipEvent_t mark(hipStream_t s) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
        if (hipEventRecord(ev, s) == hipSuccess) return ev;
        (void)hipEventDestroy(ev);
    }
    (void)hipStreamSynchronize(s);
    return nullptr;
}

void PoolBuf::release() { release_after({}); }

void PoolBuf::release_after(const std::vector<hipStream_t> &streams) {
    if (!p) return;
    std::vector<hipEvent_t> evs;
    if (hipEvent_t ev = mark(stream)) evs.push_back(ev);
    for (hipStream_t s : streams)
        if (s != stream)
            if (hipEvent_t ev = mark(s)) evs.push_back(ev);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    pool_entries().push_back({p, bytes, dev, std::move(evs)});
    p = nullptr;
    bytes = 0;
}

void pool_trim() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto &v = pool_entries();
    for (auto &e : v) {
        for (hipEvent_t ev : e.ready) {
            (void)hipEventSynchronize(ev);
            (void)hipEventDestroy(ev);
        }
        (void)hipFree(e.p);
    }
    v.clear();
}

// ------------------------------------------------------------------ kernels

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_i64(int64_t *p, size_t n, int64_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_copy_f64(double *d, const double *s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}

int fill_u32(uint32_t *p, size_t count, uint32_t v, hipStream_t s) {
    if (!count) return RLG_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, count, v);
    RLG_LAUNCHED();
    return RLG_OK;
}
int fill_i64(int64_t *p, size_t count, int64_t v, hipStream_t s) {
    if (!count) return RLG_OK;
    hipLaunchKernelGGL(k_fill_i64, dim3(ceil_div(count, 256)), dim3(256), 0, s, p, count, v);
    RLG_LAUNCHED();
    return RLG_OK;
}
int copy_f64(double *dst, const double *src, size_t count, hipStream_t s) {
    if (!count) return RLG_OK;
    hipLaunchKernelGGL(k_copy_f64, dim3(ceil_div(count, 256)), dim3(256), 0, s, dst, src, count);
    RLG_LAUNCHED();
    return RLG_OK;
}

}  // namespace rlg

using namespace rlg;

extern "C" {

const char *rlg_version(void) { return "rlgpu 0.1 (gfx950)"; }

const char *rlg_last_error_message(void) { return get_error(); }

int rlg_set_device(int device) {
    if (device < 0 || device >= 64) {
        set_error("device index %d out of range", device);
        return RLG_E_INVALID_ARG;
    }
    t_device = device;
    Device *d;
    return current_device(&d);
}

int rlg_synchronize(void) {
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_HIP(hipStreamSynchronize(d->stream));
    RLG_HIP(hipDeviceSynchronize());
    return RLG_OK;
}

int rlg_dev_alloc(size_t bytes, void **ptr) {
    if (!ptr) { set_error("null output pointer"); return RLG_E_INVALID_ARG; }
    Device *d;
    RLG_TRY(current_device(&d));
    *ptr = nullptr;
    RLG_HIP(hipMalloc(ptr, bytes ? bytes : 256));
    return RLG_OK;
}

int rlg_dev_free(void *ptr) {
    if (!ptr) return RLG_OK;
    Device *d;
    RLG_TRY(current_device(&d));
    RLG_HIP(hipFree(ptr));
    return RLG_OK;
}

int rlg_dev_upload(void *dst, const void *src, size_t bytes) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (!bytes) return RLG_OK;
    RLG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d->stream));
    RLG_HIP(hipStreamSynchronize(d->stream));
    return RLG_OK;
}

int rlg_dev_download(void *dst, const void *src, size_t bytes) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (!bytes) return RLG_OK;
    RLG_HIP(hipDeviceSynchronize());
    RLG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, d->stream));
    RLG_HIP(hipStreamSynchronize(d->stream));
    return RLG_OK;
}

int rlg_dev_copy(void *dst, const void *src, size_t bytes, void *stream) {
    Device *d;
    RLG_TRY(current_device(&d));
    if (!bytes) return RLG_OK;
    RLG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, pick_stream(d, stream)));
    return RLG_OK;
}

int rlg_prof_enable(int on) {
    auto &p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    p.on = on != 0;
    return RLG_OK;
}

int rlg_prof_reset(void) {
    auto &p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    prof_collect_locked(p);
    for (auto &x : p.launches) x = 0;
    for (auto &x : p.ms) x = 0.0;
    for (auto &x : p.work) x = 0.0;
    return RLG_OK;
}

int rlg_prof_get(const char *name, int64_t *launches, double *total_ms, double *work) {
    auto &p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    prof_collect_locked(p);
    for (size_t i = 0; i < p.names.size(); ++i) {
        if (p.names[i] == name) {
            if (launches) *launches = p.launches[i];
            if (total_ms) *total_ms = p.ms[i];
            if (work) *work = p.work[i];
            return RLG_OK;
        }
    }
    if (launches) *launches = 0;
    if (total_ms) *total_ms = 0.0;
    if (work) *work = 0.0;
    return RLG_OK;
}

int rlg_prof_dump(char *buf, size_t len) {
    auto &p = prof();
    std::lock_guard<std::mutex> lk(p.mu);
    prof_collect_locked(p);
    std::string out;
    char line[256];
    for (size_t i = 0; i < p.names.size(); ++i) {
        if (!p.launches[i]) continue;
        snprintf(line, sizeof line, "%s %lld %.6f %.6e\n", p.names[i].c_str(), (long long)p.launches[i],
                 p.ms[i], p.work[i]);
        out += line;
    }
    if (buf && len) {
        size_t n = out.size() < len - 1 ? out.size() : len - 1;
        memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return RLG_OK;
}

}  // extern "C"
