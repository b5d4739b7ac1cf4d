// Continuous-batching request loop for factor-and-solve requests.
//
// Requests enter a FIFO.  A dispatcher admits the oldest request as soon as
// (a) one of `slots` workers is idle and (b) the device bytes it needs fit
// under the memory budget together with everything in flight and the
// resident factors (least-recently-used factors that are not in use are
// evicted to make room).  Each worker owns a HIP stream, so admitted
// requests run concurrently on the device; a request that arrives while
// others are on the GPU starts as soon as a slot and memory free up.
//
// A request either carries a matrix (factored, solved, and optionally kept
// resident under factor_id) or names a resident factor (solve only).  Every
// result carries its FP64 backward error, measured on the device against
// the original matrix, and its submit-to-ready latency.
//
// Kernels are deterministic and their tiling depends only on the problem
// size, so a result is bit-identical whether its request ran alone or in a
// mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "kernels.hpp"

namespace rlg {
int solve_mixed_dev(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                    const double *b, size_t ldb, double *x, size_t ldx, const rlg_mixed_opts &o,
                    rlg_mixed_report &rep);
int backward_error_dev(hipStream_t s, size_t n, const double *a, size_t lda, size_t nrhs,
                       const double *b, size_t ldb, const double *x, size_t ldx, double *eta);
}  // namespace rlg

using namespace rlg;
using Clock = std::chrono::steady_clock;

namespace {

struct Factor {
    int64_t id = 0;
    int kind = RLG_KIND_LU;
    size_t n = 0;
    DevBuf a;       // original matrix (FP64), for residuals / refinement
    DevBuf f;       // FP64 factor
    DevBuf raw;     // LU row permutation (int64)
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

}  // namespace

struct rlg_server {
    rlg_server_opts opts{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> queue;
    std::map<int64_t, std::shared_ptr<Done>> done;
    std::map<int64_t, std::shared_ptr<Factor>> factors;
    size_t in_use = 0;       // bytes of admitted, running requests
    size_t resident = 0;     // bytes of resident factors
    int64_t next_ticket = 1;
    int idle = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;

    void evict_lru_locked(size_t need) {
        while (in_use + resident + need > opts.mem_budget_bytes) {
            std::shared_ptr<Factor> victim;
            for (auto &kv : factors)
                if (kv.second->users == 0 && (!victim || kv.second->last_use < victim->last_use))
                    victim = kv.second;
            if (!victim) return;
            resident -= victim->bytes;
            factors.erase(victim->id);
        }
    }

    // Pops the oldest admissible job; false when none is admissible now.
    bool admit_locked(Job &out) {
        if (queue.empty()) return false;
        Job &j = queue.front();
        if (j.need > opts.mem_budget_bytes) {
            // can never fit: fail it rather than block the queue
            auto d = std::make_shared<Done>();
            d->res.ticket = j.ticket;
            d->res.status = RLG_E_OOM;
            d->ready = true;
            done[j.ticket] = d;
            queue.pop_front();
            cv_done.notify_all();
            return false;
        }
        evict_lru_locked(j.need);
        if (in_use + resident + j.need > opts.mem_budget_bytes) return false;
        out = j;
        queue.pop_front();
        in_use += out.need;
        return true;
    }

    void worker(int w) {
        (void)hipSetDevice(device);
        hipStream_t s = streams[w];
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> lk(mu);
                ++idle;
                cv_work.wait(lk, [&] { return stop || admit_locked(job); });
                --idle;
                if (stop && job.ticket == 0) return;
            }
            auto d = std::make_shared<Done>();
            d->res.ticket = job.ticket;
            d->res.status = run(job, s, *d);
            d->res.latency_ms =
                std::chrono::duration<double, std::milli>(Clock::now() - job.submitted).count();
            d->res.x = d->x.empty() ? nullptr : d->x.data();
            {
                std::lock_guard<std::mutex> lk(mu);
                in_use -= job.need;
                d->ready = true;
                done[job.ticket] = d;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    int run(const Job &job, hipStream_t s, Done &d) {
        const rlg_request &r = job.req;
        const size_t n = (size_t)r.n, nr = (size_t)r.nrhs;
        std::shared_ptr<Factor> f;
        if (!r.a) {
            std::lock_guard<std::mutex> lk(mu);
            auto it = factors.find(r.factor_id);
            if (it == factors.end()) {
                set_error("factor %lld is not resident", (long long)r.factor_id);
                return RLG_E_INVALID_ARG;
            }
            f = it->second;
            if (f->n != n || f->kind != r.kind) {
                set_error("request does not match resident factor %lld", (long long)r.factor_id);
                return RLG_E_SHAPE;
            }
            f->users++;
            f->last_use = Clock::now();
        }
        // a resident factor stays pinned (users > 0) until this request ends
        struct Unpin {
            rlg_server *srv;
            std::shared_ptr<Factor> f;
            ~Unpin() {
                if (!f) return;
                std::lock_guard<std::mutex> lk(srv->mu);
                f->users--;
                f->last_use = Clock::now();
            }
        } unpin{this, r.a ? nullptr : f};
        PoolBuf b, x, tmp;
        RLG_TRY(b.alloc(nr * n * 8, s));
        RLG_TRY(x.alloc(nr * n * 8, s));
        RLG_TRY(tmp.alloc(nr * n * 8, s));
        RLG_HIP(hipMemcpyAsync(b.p, r.b, nr * n * 8, hipMemcpyHostToDevice, s));
        int st = RLG_OK;
        if (!f) {
            f = std::make_shared<Factor>();
            f->id = r.factor_id;
            f->kind = r.kind;
            f->n = n;
            f->bytes = factor_bytes(n);
            RLG_TRY(f->a.alloc(n * n * 8));
            RLG_TRY(f->f.alloc(n * n * 8));
            RLG_HIP(hipMemcpyAsync(f->a.p, r.a, n * n * 8, hipMemcpyHostToDevice, s));
            if (opts.mixed && r.kind == RLG_KIND_LU && !(r.factor_id > 0)) {
                // one-shot request: refined mixed-precision solve
                rlg_mixed_opts o;
                rlg_mixed_opts_default(&o);
                o.tol = opts.tol;
                rlg_mixed_report rep;
                st = solve_mixed_dev(s, n, f->a.as<double>(), n, nr, b.as<double>(), n, x.as<double>(),
                                     n, o, rep);
                d.res.fell_back = rep.fell_back;
                d.res.backward_error = rep.backward_error;
                f.reset();
            } else {
                RLG_HIP(hipMemcpyAsync(f->f.p, f->a.p, n * n * 8, hipMemcpyDeviceToDevice, s));
                if (r.kind == RLG_KIND_LU) {
                    PoolBuf ipiv, info, work;
                    RLG_TRY(ipiv.alloc(n * 8, s));
                    RLG_TRY(info.alloc(8, s));
                    RLG_TRY(work.alloc(lu_work_bytes(n, 8), s));
                    RLG_TRY(lu_factor<double>(s, n, f->f.as<double>(), n, ipiv.as<int64_t>(),
                                              info.as<int64_t>(), work.p));
                    int64_t inf = 0;
                    int timeout = 0;
                    std::vector<int64_t> ip(n), raw(n);
                    RLG_HIP(hipMemcpyAsync(&inf, info.p, 8, hipMemcpyDeviceToHost, s));
                    RLG_HIP(hipMemcpyAsync(&timeout, lu_timeout_flag(work.p, n, 8), 4, hipMemcpyDeviceToHost, s));
                    RLG_HIP(hipMemcpyAsync(ip.data(), ipiv.p, n * 8, hipMemcpyDeviceToHost, s));
                    RLG_HIP(hipStreamSynchronize(s));
                    if (timeout) { set_error("LU panel barrier timed out"); return RLG_E_DEVICE; }
                    if (inf < (int64_t)n) { set_error("singular matrix"); return RLG_E_DIV_BY_ZERO; }
                    perm_from_ipiv(n, ip.data(), raw.data());
                    RLG_TRY(f->raw.alloc(n * 8));
                    RLG_HIP(hipMemcpyAsync(f->raw.p, raw.data(), n * 8, hipMemcpyHostToDevice, s));
                    RLG_HIP(hipStreamSynchronize(s));
                } else {
                    PoolBuf info, work;
                    RLG_TRY(info.alloc(8, s));
                    RLG_TRY(work.alloc(chol_work_bytes(n, 8), s));
                    RLG_TRY(chol_factor<double>(s, n, f->f.as<double>(), n, info.as<int64_t>(), work.p));
                    int64_t inf = 0;
                    RLG_HIP(hipMemcpyAsync(&inf, info.p, 8, hipMemcpyDeviceToHost, s));
                    RLG_HIP(hipStreamSynchronize(s));
                    if (inf < (int64_t)n) { set_error("not positive definite"); return RLG_E_DECOMP_FAILURE; }
                }
            }
        }
        if (f) {
            if (f->kind == RLG_KIND_LU)
                RLG_TRY(lu_solve<double>(s, n, f->f.as<double>(), n, f->raw.as<int64_t>(), nr,
                                         b.as<double>(), n, x.as<double>(), n, tmp.as<double>()));
            else
                RLG_TRY(chol_solve<double>(s, n, f->f.as<double>(), n, nr, b.as<double>(), n,
                                           x.as<double>(), n));
            double eta = 0.0;
            RLG_TRY(backward_error_dev(s, n, f->a.as<double>(), n, nr, b.as<double>(), n,
                                       x.as<double>(), n, &eta));
            d.res.backward_error = eta;
            std::lock_guard<std::mutex> lk(mu);
            if (r.a && r.factor_id > 0) {
                f->last_use = Clock::now();
                auto old = factors.find(r.factor_id);
                if (old != factors.end()) resident -= old->second->bytes;
                factors[r.factor_id] = f;
                resident += f->bytes;
            }
        }
        if (st != RLG_OK) return st;
        d.x.resize(nr * n);
        RLG_HIP(hipMemcpyAsync(d.x.data(), x.p, nr * n * 8, hipMemcpyDeviceToHost, s));
        RLG_HIP(hipStreamSynchronize(s));
        return RLG_OK;
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

int rlg_server_create(const rlg_server_opts *o, rlg_server **out) {
    if (!out) { set_error("null output"); return RLG_E_INVALID_ARG; }
    *out = nullptr;
    Device *d;
    RLG_TRY(current_device(&d));
    auto *s = new rlg_server();
    if (o) s->opts = *o; else rlg_server_opts_default(&s->opts);
    if (s->opts.slots < 1) s->opts.slots = 1;
    if (s->opts.slots > 16) s->opts.slots = 16;
    if (!(s->opts.tol > 0)) s->opts.tol = 1e-13;
    s->device = d->id;
    for (int i = 0; i < s->opts.slots; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            set_error("stream creation failed");
            for (auto x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return RLG_E_DEVICE;
        }
        s->streams.push_back(st);
    }
    for (int i = 0; i < s->opts.slots; ++i) s->workers.emplace_back(&rlg_server::worker, s, i);
    *out = s;
    return RLG_OK;
}

int rlg_server_submit(rlg_server *s, const rlg_request *r, int64_t *ticket) {
    if (!s || !r || !ticket) { set_error("null argument"); return RLG_E_INVALID_ARG; }
    if (r->n < 1 || r->nrhs < 1 || !r->b || (r->kind != RLG_KIND_LU && r->kind != RLG_KIND_CHOL) ||
        (!r->a && r->factor_id <= 0)) {
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
        j.ticket = s->next_ticket++;
        s->queue.push_back(j);
    }
    *ticket = j.ticket;
    s->cv_work.notify_one();
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
    s->factors.erase(it);
    return RLG_OK;
}

void rlg_server_destroy(rlg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete s;
}

}  // extern "C"
