// Continuous-batching request loop for factor-and-solve requests.
//
// One host thread (the dispatcher) owns every HIP call of the server.  It
// keeps up to `slots` requests in flight, each on its own stream, and
// advances them as a state machine:
//
//   ADMIT   the oldest queued request is admitted as soon as a slot is free
//           and the device bytes it needs fit under the memory budget with
//           everything in flight and the resident factors (idle resident
//           factors are evicted least-recently-used first to make room);
//   FACTOR  copy-in through the slot's pinned staging buffer, factorization
//           kernels enqueued, an event recorded;
//   SOLVE   once the event has completed (polled, never waited on), the
//           pivots are checked and the triangular solves and the FP64
//           backward error are enqueued, then the copy-out;
//   DONE    the result (solution, backward error, latency) is posted.
//
// Requests that arrive while others are on the GPU are admitted as slots and
// memory free up, and their kernels overlap on the device through the
// slots' streams.  Device buffers come from the library's pool and resident
// factors are freed only when evicted and idle, so no hipMalloc or hipFree
// runs while another slot's work is in flight once the pool is warm.
//
// Kernels are deterministic and their tiling depends only on the problem
// size, so a result is bit-identical whether its request ran alone or in a
// mix.
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "kernels.hpp"

namespace rlg {
int solve_mixed_dev(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                    const double *b, size_t ldb, double *x, size_t ldx, const rlg_mixed_opts &o,
                    rlg_mixed_report &rep);
int backward_error_enqueue(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                           const double *b, size_t ldb, const double *x, size_t ldx, double *norms,
                           double *res, double *rs);
double backward_error_from(const double *norms_host, size_t nrhs);
}  // namespace rlg

using namespace rlg;
using Clock = std::chrono::steady_clock;

namespace {

struct Factor {
    int64_t id = 0;
    int kind = RLG_KIND_LU;
    size_t n = 0;
    PoolBuf a;       // original matrix (FP64), for residuals
    PoolBuf f;       // FP64 factor
    PoolBuf raw;     // LU row permutation (int64)
    size_t bytes = 0;
    int users = 0;
    Clock::time_point last_use;
};

struct Job {
    int64_t ticket = 0;
    rlg_request req{};
    Clock::time_point submitted;
    size_t need = 0;
};

struct Done {
    rlg_result res{};
    std::vector<double> x;
    bool ready = false;
};

size_t factor_bytes(size_t n) { return n * n * 8 * 2 + n * 8; }

size_t request_bytes(const rlg_request &r, bool mixed) {
    const size_t n = (size_t)r.n, nr = (size_t)r.nrhs;
    size_t b = 4 * nr * n * 8 + lu_work_bytes(n, 8) + (1 << 20);
    if (r.a) {
        b += factor_bytes(n);
        if (mixed) b += n * n * 4 + 4 * nr * n * 4;
    }
    return b;
}

enum Stage { IDLE = 0, FACTORED, SOLVED };

// One in-flight request and the per-slot resources it uses.
struct Slot {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    double *pinned = nullptr;      // staging for copy-in / copy-out
    size_t pinned_elems = 0;
    Stage stage = IDLE;
    Job job;
    std::shared_ptr<Done> done;
    std::shared_ptr<Factor> f;
    PoolBuf b, x, tmp, ipiv, info, work, norms, res, rs;
    int64_t *ipiv_h = nullptr;     // pinned
    int64_t *info_h = nullptr;     // pinned (info, timeout)
    double *norms_h = nullptr;     // pinned (3 * nrhs + 1)
    bool mixed_path = false;
    rlg_mixed_report rep{};
};

}  // namespace

struct rlg_server {
    rlg_server_opts opts{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> queue;
    std::map<int64_t, std::shared_ptr<Done>> done;
    std::map<int64_t, std::shared_ptr<Factor>> factors;
    std::vector<std::shared_ptr<Factor>> graveyard;  // evicted, freed by the dispatcher
    size_t in_use = 0;
    size_t resident = 0;
    int64_t next_ticket = 1;
    bool stop = false;
    std::thread dispatcher;
    std::vector<std::unique_ptr<Slot>> slots;

    // ---------------------------------------------------------- admission
    void evict_lru_locked(size_t need) {
        while (in_use + resident + need > opts.mem_budget_bytes) {
            std::shared_ptr<Factor> victim;
            for (auto &kv : factors)
                if (kv.second->users == 0 && (!victim || kv.second->last_use < victim->last_use))
                    victim = kv.second;
            if (!victim) return;
            resident -= victim->bytes;
            graveyard.push_back(victim);
            factors.erase(victim->id);
        }
    }

    bool admit_locked(Job &out) {
        while (!queue.empty()) {
            Job &j = queue.front();
            if (j.need > opts.mem_budget_bytes) {  // can never fit: fail it, keep going
                post_locked(j.ticket, RLG_E_OOM, nullptr);
                queue.pop_front();
                continue;
            }
            evict_lru_locked(j.need);
            if (in_use + resident + j.need > opts.mem_budget_bytes) return false;
            out = j;
            queue.pop_front();
            in_use += out.need;
            return true;
        }
        return false;
    }

    void post_locked(int64_t ticket, int status, std::shared_ptr<Done> d) {
        if (!d) d = std::make_shared<Done>();
        d->res.ticket = ticket;
        d->res.status = status;
        d->res.x = d->x.empty() ? nullptr : d->x.data();
        d->ready = true;
        done[ticket] = d;
        cv_done.notify_all();
    }

    // ---------------------------------------------------------- stages
    int ensure_pinned(Slot &sl, size_t elems) {
        if (sl.pinned_elems >= elems) return RLG_OK;
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        sl.pinned = nullptr;
        sl.pinned_elems = 0;
        RLG_HIP(hipHostMalloc((void **)&sl.pinned, elems * 8, hipHostMallocDefault));
        sl.pinned_elems = elems;
        return RLG_OK;
    }

    // copy host -> device through the slot's pinned buffer (synchronous for the slot)
    int stage_in(Slot &sl, void *dst, const double *src, size_t elems) {
        RLG_TRY(ensure_pinned(sl, elems));
        std::memcpy(sl.pinned, src, elems * 8);
        RLG_HIP(hipMemcpyAsync(dst, sl.pinned, elems * 8, hipMemcpyHostToDevice, sl.s));
        RLG_HIP(hipEventRecord(sl.ev, sl.s));
        RLG_HIP(hipEventSynchronize(sl.ev));  // the staging buffer is reused next
        return RLG_OK;
    }

    int start(Slot &sl) {
        const rlg_request &r = sl.job.req;
        const size_t n = (size_t)r.n, nr = (size_t)r.nrhs;
        hipStream_t s = sl.s;
        sl.done = std::make_shared<Done>();
        sl.mixed_path = false;
        sl.f.reset();
        if (!r.a) {
            std::lock_guard<std::mutex> lk(mu);
            auto it = factors.find(r.factor_id);
            if (it == factors.end()) {
                set_error("factor %lld is not resident", (long long)r.factor_id);
                return RLG_E_INVALID_ARG;
            }
            if (it->second->n != n || it->second->kind != r.kind) {
                set_error("request does not match resident factor %lld", (long long)r.factor_id);
                return RLG_E_SHAPE;
            }
            sl.f = it->second;
            sl.f->users++;
            sl.f->last_use = Clock::now();
        }
        RLG_TRY(sl.b.alloc(nr * n * 8, s));
        RLG_TRY(sl.x.alloc(nr * n * 8, s));
        RLG_TRY(sl.tmp.alloc(nr * n * 8, s));
        RLG_TRY(sl.norms.alloc((3 * nr + 1) * 8, s));
        RLG_TRY(sl.res.alloc(nr * n * 8, s));
        RLG_TRY(sl.rs.alloc(n * 8, s));
        RLG_TRY(stage_in(sl, sl.b.p, r.b, nr * n));
        if (sl.f) {  // resident factor: straight to the solve
            sl.stage = FACTORED;
            return RLG_OK;
        }
        auto f = std::make_shared<Factor>();
        f->id = r.factor_id;
        f->kind = r.kind;
        f->n = n;
        f->bytes = factor_bytes(n);
        RLG_TRY(f->a.alloc(n * n * 8, s));
        RLG_TRY(stage_in(sl, f->a.p, r.a, n * n));
        sl.f = f;
        if (opts.mixed && r.kind == RLG_KIND_LU && !(r.factor_id > 0)) {
            // one-shot refined mixed-precision solve; it synchronizes its stream
            rlg_mixed_opts o;
            rlg_mixed_opts_default(&o);
            o.tol = opts.tol;
            sl.mixed_path = true;
            RLG_TRY(solve_mixed_dev(s, n, f->a.as<double>(), n, nr, sl.b.as<double>(), n,
                                    sl.x.as<double>(), n, o, sl.rep));
            sl.stage = FACTORED;
            return RLG_OK;
        }
        RLG_TRY(f->f.alloc(n * n * 8, s));
        RLG_HIP(hipMemcpyAsync(f->f.p, f->a.p, n * n * 8, hipMemcpyDeviceToDevice, s));
        if (r.kind == RLG_KIND_LU) {
            RLG_TRY(sl.ipiv.alloc(n * 8, s));
            RLG_TRY(sl.info.alloc(8, s));
            RLG_TRY(sl.work.alloc(lu_work_bytes(n, 8), s));
            RLG_TRY(lu_factor<double>(s, n, f->f.as<double>(), n, sl.ipiv.as<int64_t>(),
                                      sl.info.as<int64_t>(), sl.work.p));
            RLG_HIP(hipMemcpyAsync(sl.ipiv_h, sl.ipiv.p, n * 8, hipMemcpyDeviceToHost, s));
            RLG_HIP(hipMemcpyAsync(sl.info_h, sl.info.p, 8, hipMemcpyDeviceToHost, s));
            RLG_HIP(hipMemcpyAsync(sl.info_h + 1, lu_timeout_flag(sl.work.p, n, 8), 4,
                                   hipMemcpyDeviceToHost, s));
        } else {
            RLG_TRY(sl.info.alloc(8, s));
            RLG_TRY(sl.work.alloc(chol_work_bytes(n, 8), s));
            RLG_TRY(chol_factor<double>(s, n, f->f.as<double>(), n, sl.info.as<int64_t>(), sl.work.p));
            RLG_HIP(hipMemcpyAsync(sl.info_h, sl.info.p, 8, hipMemcpyDeviceToHost, s));
            sl.info_h[1] = 0;
        }
        RLG_HIP(hipEventRecord(sl.ev, s));
        sl.stage = FACTORED;
        return RLG_OK;
    }

    // called once the FACTORED event has completed
    int solve(Slot &sl) {
        const rlg_request &r = sl.job.req;
        const size_t n = (size_t)r.n, nr = (size_t)r.nrhs;
        hipStream_t s = sl.s;
        Factor &f = *sl.f;
        if (!sl.mixed_path) {
            if (r.a) {  // fresh factorization: check it and finish the pivots
                if (sl.info_h[1]) { set_error("LU panel barrier timed out"); return RLG_E_DEVICE; }
                if (sl.info_h[0] < (int64_t)n) {
                    set_error(r.kind == RLG_KIND_LU ? "singular matrix" : "not positive definite");
                    return r.kind == RLG_KIND_LU ? RLG_E_DIV_BY_ZERO : RLG_E_DECOMP_FAILURE;
                }
                if (r.kind == RLG_KIND_LU) {
                    std::vector<int64_t> raw(n);
                    perm_from_ipiv(n, sl.ipiv_h, raw.data());
                    RLG_TRY(f.raw.alloc(n * 8, s));
                    RLG_TRY(stage_in(sl, f.raw.p, reinterpret_cast<const double *>(raw.data()), n));
                }
            }
            if (f.kind == RLG_KIND_LU)
                RLG_TRY(lu_solve<double>(s, n, f.f.as<double>(), n, f.raw.as<int64_t>(), nr,
                                         sl.b.as<double>(), n, sl.x.as<double>(), n, sl.tmp.as<double>()));
            else
                RLG_TRY(chol_solve<double>(s, n, f.f.as<double>(), n, nr, sl.b.as<double>(), n,
                                           sl.x.as<double>(), n, sl.tmp.as<double>()));
            RLG_TRY(backward_error_enqueue(s, n, f.a.as<double>(), n, nr, sl.b.as<double>(), n,
                                           sl.x.as<double>(), n, sl.norms.as<double>(),
                                           sl.res.as<double>(), sl.rs.as<double>()));
            RLG_HIP(hipMemcpyAsync(sl.norms_h, sl.norms.p, (3 * nr + 1) * 8, hipMemcpyDeviceToHost, s));
        }
        RLG_TRY(ensure_pinned(sl, nr * n));
        RLG_HIP(hipMemcpyAsync(sl.pinned, sl.x.p, nr * n * 8, hipMemcpyDeviceToHost, s));
        RLG_HIP(hipEventRecord(sl.ev, s));
        sl.stage = SOLVED;
        return RLG_OK;
    }

    // called once the SOLVED event has completed
    void finish(Slot &sl, int status) {
        const rlg_request &r = sl.job.req;
        auto d = sl.done ? sl.done : std::make_shared<Done>();
        if (status == RLG_OK) {
            const size_t count = (size_t)r.n * (size_t)r.nrhs;
            d->x.assign(sl.pinned, sl.pinned + count);
            if (sl.mixed_path) {
                d->res.fell_back = sl.rep.fell_back;
                d->res.backward_error = sl.rep.backward_error;
            } else {
                d->res.backward_error = backward_error_from(sl.norms_h, (size_t)r.nrhs);
            }
        }
        d->res.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - sl.job.submitted).count();
        std::shared_ptr<Factor> keep = sl.f;
        sl.f.reset();
        sl.done.reset();
        sl.stage = IDLE;
        for (PoolBuf *p : {&sl.b, &sl.x, &sl.tmp, &sl.ipiv, &sl.info, &sl.work, &sl.norms, &sl.res, &sl.rs})
            p->release();
        std::lock_guard<std::mutex> lk(mu);
        in_use -= sl.job.need;
        if (keep) {
            keep->last_use = Clock::now();
            if (!r.a) {
                keep->users--;
            } else if (status == RLG_OK && r.factor_id > 0 && !sl.mixed_path) {
                auto old = factors.find(r.factor_id);
                if (old != factors.end()) {
                    resident -= old->second->bytes;
                    if (old->second->users == 0) graveyard.push_back(old->second);
                }
                factors[r.factor_id] = keep;
                resident += keep->bytes;
            } else {
                graveyard.push_back(keep);  // freed below, outside any in-flight use
            }
        }
        post_locked(sl.job.ticket, status, d);
    }

    void loop() {
        (void)hipSetDevice(device);
        for (;;) {
            bool busy = false;
            // admit into idle slots
            for (auto &sp : slots) {
                Slot &sl = *sp;
                if (sl.stage != IDLE) { busy = true; continue; }
                Job job;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (!admit_locked(job)) continue;
                }
                sl.job = job;
                busy = true;
                const int st = start(sl);
                if (st != RLG_OK) finish(sl, st);
            }
            // advance in-flight slots whose event has completed
            for (auto &sp : slots) {
                Slot &sl = *sp;
                if (sl.stage == IDLE) continue;
                if (hipEventQuery(sl.ev) == hipErrorNotReady) continue;
                if (sl.stage == FACTORED) {
                    const int st = solve(sl);
                    if (st != RLG_OK) finish(sl, st);
                } else {
                    finish(sl, RLG_OK);
                }
            }
            {
                std::unique_lock<std::mutex> lk(mu);
                // evicted factors are idle: release them to the pool now
                graveyard.clear();
                bool inflight = false;
                for (auto &sp : slots) inflight |= sp->stage != IDLE;
                if (stop && queue.empty() && !inflight) return;
                if (!inflight && queue.empty())
                    cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            }
            if (busy) std::this_thread::yield();
        }
    }
};

extern "C" {

void rlg_server_opts_default(rlg_server_opts *o) {
    if (!o) return;
    o->mem_budget_bytes = (size_t)64 << 30;
    o->slots = 4;
    o->mixed = 0;
    o->tol = 1e-13;
}

// The request loop stays switched off until it has run on the GPU; the
// environment variable lets its own tests reach it.
static bool server_enabled() {
    const char *e = getenv("RLG_ENABLE_SERVER");
    return e && e[0] == '1';
}

int rlg_server_create(const rlg_server_opts *o, rlg_server **out) {
    if (!out) { set_error("null output"); return RLG_E_INVALID_ARG; }
    *out = nullptr;
    if (!server_enabled()) {
        set_error("rlg_server_* is disabled in this build: an earlier form of this request loop "
                  "stalled on the GPU and this form has not been validated there; use the "
                  "per-system and batched entry points (or rulinalg_amd.serving)");
        return RLG_E_INVALID_ARG;
    }
    Device *d;
    RLG_TRY(current_device(&d));
    auto *s = new rlg_server();
    if (o) s->opts = *o; else rlg_server_opts_default(&s->opts);
    if (s->opts.slots < 1) s->opts.slots = 1;
    if (s->opts.slots > 8) s->opts.slots = 8;
    if (!(s->opts.tol > 0)) s->opts.tol = 1e-13;
    s->device = d->id;
    for (int i = 0; i < s->opts.slots; ++i) s->slots.emplace_back(new Slot());
    int st = RLG_OK;
    for (auto &sp : s->slots) {
        Slot &sl = *sp;
        if (hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&sl.ipiv_h, 16384 * 8, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&sl.info_h, 16, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&sl.norms_h, (3 * RLG_BATCH_MAX_NRHS + 1) * 8 * 4,
                          hipHostMallocDefault) != hipSuccess) {
            set_error("server slot creation failed");
            st = RLG_E_DEVICE;
            break;
        }
    }
    if (st != RLG_OK) {
        rlg_server_destroy(s);
        return st;
    }
    s->dispatcher = std::thread(&rlg_server::loop, s);
    *out = s;
    return RLG_OK;
}

int rlg_server_submit(rlg_server *s, const rlg_request *r, int64_t *ticket) {
    if (!s || !r || !ticket) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    if (r->n < 1 || r->nrhs < 1 || r->nrhs > 4 * RLG_BATCH_MAX_NRHS || !r->b ||
        (r->kind != RLG_KIND_LU && r->kind != RLG_KIND_CHOL) || (!r->a && r->factor_id <= 0)) {
        set_error("invalid request (kind %d, n %d, nrhs %d)", r->kind, r->n, r->nrhs);
        return RLG_E_INVALID_ARG;
    }
    if (r->n > 16384) { set_error("request order %d above 16384", r->n); return RLG_E_INVALID_ARG; }
    Job j;
    j.req = *r;
    j.submitted = Clock::now();
    j.need = request_bytes(*r, s->opts.mixed != 0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) { set_error("server is shutting down"); return RLG_E_INVALID_ARG; }
        j.ticket = s->next_ticket++;
        s->queue.push_back(j);
    }
    *ticket = j.ticket;
    s->cv_work.notify_all();
    return RLG_OK;
}

int rlg_server_wait(rlg_server *s, int64_t ticket, rlg_result *out) {
    if (!s || !out) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket <= 0 || ticket >= s->next_ticket) { set_error("unknown ticket"); return RLG_E_INVALID_ARG; }
    s->cv_done.wait(lk, [&] {
        auto it = s->done.find(ticket);
        return it != s->done.end() && it->second->ready;
    });
    *out = s->done[ticket]->res;
    return RLG_OK;
}

int rlg_server_release(rlg_server *s, int64_t ticket) {
    if (!s) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    s->done.erase(ticket);
    return RLG_OK;
}

int rlg_server_evict(rlg_server *s, int64_t factor_id) {
    if (!s) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->factors.find(factor_id);
    if (it == s->factors.end()) return RLG_OK;
    if (it->second->users) { set_error("factor in use"); return RLG_E_INVALID_ARG; }
    s->resident -= it->second->bytes;
    s->graveyard.push_back(it->second);
    s->factors.erase(it);
    s->cv_work.notify_all();
    return RLG_OK;
}

void rlg_server_destroy(rlg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    s->factors.clear();
    s->graveyard.clear();
    for (auto &sp : s->slots) {
        Slot &sl = *sp;
        if (sl.s) (void)hipStreamSynchronize(sl.s);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.ipiv_h) (void)hipHostFree(sl.ipiv_h);
        if (sl.info_h) (void)hipHostFree(sl.info_h);
        if (sl.norms_h) (void)hipHostFree(sl.norms_h);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.s) (void)hipStreamDestroy(sl.s);
    }
    delete s;
}

}  // extern "C"
