// Shared runtime pieces of librlgpu: error reporting, device context,
// streams, workspace allocation and the per-kernel event profiler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rlgpu.h"

namespace rlg {

// ------------------------------------------------------------------ errors

void set_error(const char *fmt, ...);
const char *get_error();

#define RLG_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::rlg::set_error("%s failed: %s (%s:%d)", #call,                   \
                             hipGetErrorString(e_), __FILE__, __LINE__);       \
            return e_ == hipErrorOutOfMemory ? RLG_E_OOM : RLG_E_DEVICE;       \
        }                                                                      \
    } while (0)

#define RLG_TRY(call)                                                          \
    do {                                                                       \
        int s_ = (call);                                                       \
        if (s_ != RLG_OK) return s_;                                           \
    } while (0)

// ------------------------------------------------------------------ device

struct Device {
    int id = 0;
    hipStream_t stream = nullptr;    // the library's default stream
    int num_cus = 256;
    size_t lds_per_block = 65536;
};

// Per-thread current device (rlg_set_device); the Device objects are
// created lazily and live until process exit.
int current_device(Device **dev);
hipStream_t pick_stream(Device *dev, void *user_stream);
// A second stream of the current device (default priority) for work that
// overlaps the caller's stream, ordered by events.  It lives for one call.
// With two such streams kept per thread after the mixed LU's look-ahead had
// used them, the FP64 LU at order 16384 (cooperative panels) measured
// 196 ms instead of 142 ms; created per call, it does not.  Other streams,
// idle or after running a kernel, did not slow it (measured at order 4096),
// so the cause lies in how the kept streams had been used.
struct AuxStream {
    hipStream_t s = nullptr;
    int create();
    ~AuxStream();
    AuxStream() = default;
    AuxStream(const AuxStream &) = delete;
    AuxStream &operator=(const AuxStream &) = delete;
};

// ------------------------------------------------------------------ memory

// RAII device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t n) {
        release();
        if (n == 0) n = 256;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? RLG_E_OOM : RLG_E_DEVICE;
        }
        bytes = n;
        return RLG_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Pooled device workspace.  A released buffer goes back to a per-device
// free list together with an event recorded on the stream that used it; a
// later acquire of a buffer that is at least as large (and at most twice as
// large) reuses it after making the acquiring stream wait for that event.
// Steady-state calls of the solvers therefore do no hipMalloc.
struct PoolBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    int dev = 0;
    PoolBuf() = default;
    PoolBuf(const PoolBuf &) = delete;
    PoolBuf &operator=(const PoolBuf &) = delete;
    ~PoolBuf() { release(); }
    int alloc(size_t n, hipStream_t s);
    void release();
    // Back to the pool once every stream in `streams` (and the buffer's own
    // stream) has reached the point where this is called; nothing waits on
    // the host.
    void release_after(const std::vector<hipStream_t> &streams);
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// The streams that ran work reading a factor handle's device memory, so that
// freeing the handle waits for exactly those streams.
struct StreamUse {
    std::mutex mu;
    std::vector<hipStream_t> streams;
    void add(hipStream_t s) {
        std::lock_guard<std::mutex> lk(mu);
        for (auto x : streams)
            if (x == s) return;
        streams.push_back(s);
    }
    std::vector<hipStream_t> all() {
        std::lock_guard<std::mutex> lk(mu);
        return streams;
    }
};
// Frees every pooled buffer that is not in use (all devices).
void pool_trim();

// ------------------------------------------------------------------ profiler

// When enabled, every launch made through Prof is bracketed by a pair of
// HIP events on its own stream; times are collected lazily at query time.
struct ProfScope {
    int cls;
    hipStream_t s;
    double work;   // algorithmic flops (or bytes) of the bracketed launches
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(int cls, hipStream_t s, double work = 0.0);
    ~ProfScope();
};
int prof_class(const char *name);
bool prof_on();

#define RLG_CAT2(a, b) a##b
#define RLG_CAT(a, b) RLG_CAT2(a, b)
#define RLG_PROF(name, stream)                                                 \
    static const int RLG_CAT(rlg_prof_cls_, __LINE__) = ::rlg::prof_class(name); \
    ::rlg::ProfScope RLG_CAT(rlg_prof_scope_, __LINE__)(RLG_CAT(rlg_prof_cls_, __LINE__), stream)
#define RLG_PROF_W(name, stream, work)                                         \
    static const int RLG_CAT(rlg_prof_cls_, __LINE__) = ::rlg::prof_class(name); \
    ::rlg::ProfScope RLG_CAT(rlg_prof_scope_, __LINE__)(RLG_CAT(rlg_prof_cls_, __LINE__), stream, work)

// Checks for a launch failure right after a kernel launch.
#define RLG_LAUNCHED()                                                         \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            ::rlg::set_error("kernel launch failed: %s (%s:%d)",               \
                             hipGetErrorString(e_), __FILE__, __LINE__);       \
            return RLG_E_DEVICE;                                               \
        }                                                                      \
    } while (0)

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Fill `count` elements of a device array with a value (used in place of
// hipMemsetAsync so that the work can live inside a captured graph).
int fill_u32(uint32_t *p, size_t count, uint32_t v, hipStream_t s);
int fill_i64(int64_t *p, size_t count, int64_t v, hipStream_t s);
int copy_f64(double *dst, const double *src, size_t count, hipStream_t s);

}  // namespace rlg
