"""Continuous-batching request loop over the batched and per-system solvers.

A dispatcher (the thread that calls ``run``) admits requests in arrival order
as they arrive, while earlier ones are still on the GPU: a request is
admitted when its device bytes fit under the memory budget together with the
bytes of the requests in flight and of the resident factors (the
least-recently-used resident factor that no admitted request needs is
evicted to make room; a request that can never fit fails with RLG_E_OOM).
Admitted requests go to workers, each a host thread with its own HIP stream:

* one-shot systems of order <= 512 collect in a queue that the batch worker
  empties whenever it is idle: all of them run in ONE persistent batched
  launch (``rlg_batch_solve_f64_dev``), so their work overlaps on the device,
  and systems arriving meanwhile form the next launch.  Mixed mode uses the
  same FP64 launch for them unless ``small_mixed`` asks for
  ``rlg_batch_solve_mixed_dev``: at these orders the FP32 factor gains
  nothing (the FP32 16x16x4 MFMA runs at the FP64 rate) and its refinement
  doubles the launch's time.  Measured on the 2,000 req/s mixed trace, the
  FP64 launch halved the small systems' median service time (about 4 against
  9 ms) at the same sustained throughput;
* larger systems, and requests that keep or reuse a resident factor, go to
  the first free slot worker, which runs the per-system path
  (``rlg_{lu,chol}_decompose_f64_dev`` + ``_solve_f64_dev``, or
  ``rlg_solve_mixed_f64_dev`` for one-shot mixed LU) on its stream; the slot
  workers' systems run on the device at the same time.  In mixed mode the
  large systems' low-precision factorizations use tournament pivoting
  (``mixed_pivot``, default ``PIVOT_TOURNAMENT``): its panels combine their
  workgroups' candidates without waiting for one another, so they are
  ordinary launches that run beside the other slots' work, where the
  partial-pivoting panel is a cooperative launch that the runtime does not
  overlap with other streams' kernels.  The FP64 refinement and fallback
  make the pivot choice invisible in the result's accuracy.  The plain mode
  keeps the reference's partial pivoting.

Every call a worker makes is ordered on its own stream and a worker waits
only for its own stream, so no request waits for another's work to finish;
a request's latency ends when its worker's stream has delivered the
solution.  A request that keeps its factor (``keep_id``) leaves the handle
resident under that id, with its original matrix for the backward error; a
later request with ``reuse_id`` solves new right-hand sides against it
without refactoring (several such solves may run at once: the handle is
immutable).  A reuse request that arrives while the factor it names is
still being computed waits for it.  Reuse solves take a free slot worker
before any queued factorization: each is two triangular solves, far shorter
than the factorizations ahead of it.

The reference has no serving layer (rulinalg is a library of synchronous
calls); the loop serves its calls: PartialPivLu::decompose + solve
(src/matrix/decomposition/lu.rs:163-244, solves on `&self` at lu.rs:231) and
Cholesky::decompose + solve (src/matrix/decomposition/cholesky.rs:116-202,
cholesky.rs:194).  Results do not depend on what else is in flight: the
batched kernel solves every system in one workgroup with a fixed operation
order, and the per-system path is deterministic.
"""
import collections
import ctypes as C
import queue
import threading
import time

from . import _lib

KIND_LU, KIND_CHOL = _lib.KIND_LU, _lib.KIND_CHOL
SMALL_MAX = 512  # RLG_BATCH_MAX_N
RLG_E_OOM = 9


class Request:
    """A solve request.  ``a`` (n*n) and ``b`` (nrhs*n) are device tensors
    holding the input as it arrived (``a`` may be None for a ``reuse_id``
    request); ``x`` (nrhs*n) receives the solution."""

    __slots__ = ("rid", "kind", "n", "nrhs", "a", "b", "x", "arrival", "keep_id", "reuse_id",
                 "check", "status", "done", "eta", "path", "fell_back", "t_arr", "t_start", "need", "_kept")

    def __init__(self, rid, kind, n, nrhs, a, b, x, arrival, keep_id=None, reuse_id=None, check=True):
        self.rid, self.kind, self.n, self.nrhs = rid, kind, n, nrhs
        self.a, self.b, self.x, self.arrival = a, b, x, arrival
        self.keep_id, self.reuse_id, self.check = keep_id, reuse_id, check
        self.status, self.done, self.eta, self.path, self.fell_back = None, None, None, None, 0
        self.t_arr = None  # arrival on the host clock, set by ServingLoop.run
        self.t_start = None  # when a worker started on it
        self.need = 0      # device bytes held while in flight
        self._kept = None  # (handle, factor tensor) of a kept factor, for the dispatcher

    @property
    def latency_ms(self):
        return None if self.done is None else (self.done - self.t_arr) * 1e3


class _Worker(threading.Thread):
    """A host thread with its own stream that runs the jobs it is given, one
    at a time, and reports each finished job on the completion queue."""

    def __init__(self, loop, name):
        super().__init__(name=name, daemon=True)
        self.loop = loop
        self.jobs = queue.Queue()
        self.stream = loop._new_stream()
        self.busy = False

    def run(self):
        while True:
            job = self.jobs.get()
            if job is None:
                return
            fn, arg = job
            t = time.perf_counter()
            for r in (arg if isinstance(arg, list) else [arg[0]]):
                r.t_start = t
            try:
                fn(self, arg)
                err = None
            except Exception as e:  # reported to the dispatcher, which re-raises it
                err = e
            self.loop.completions.put((self, arg, err))


class ServingLoop:
    def __init__(self, lib, torch, dev, mem_budget_bytes=64 << 30, mixed=False, tol=1e-13, slots=4,
                 mixed_pivot=_lib.PIVOT_TOURNAMENT, small_mixed=False):
        self.lib, self.torch, self.dev = lib, torch, dev
        self.budget = mem_budget_bytes
        self.mixed = mixed
        self.small_mixed = mixed and small_mixed  # the mixed batched launch for orders <= 512
        self.mixed_pivot = mixed_pivot
        self.tol = tol
        self.resident = collections.OrderedDict()  # id -> (handle, kind, n, work, a, bytes)
        self.resident_bytes = 0
        self.completions = queue.Queue()
        self.batch_worker = _Worker(self, "rlg-batch")
        self.slot_workers = [_Worker(self, f"rlg-slot{i}") for i in range(max(1, slots))]
        for w in [self.batch_worker] + self.slot_workers:
            w.start()

    # ---------------------------------------------------------------- streams
    def _new_stream(self):
        return self.torch.cuda.Stream(device=self.dev)

    def _sptr(self, worker):
        return C.c_void_p(worker.stream.cuda_stream)

    # ---------------------------------------------------------------- memory
    @staticmethod
    def _need(r):
        """Device bytes a request holds while it runs (the factor copy, or the
        batch's copies of A and B), plus what stays resident if kept."""
        if r.reuse_id is not None:
            return 0
        return 8 * r.n * r.n + 8 * r.nrhs * r.n + (16 * r.n * r.n if r.keep_id is not None else 0)

    def _evict_for(self, need, protected):
        """Evicts least-recently-used resident factors for which
        `protected(id)` is false until `need` more bytes fit."""
        while self.resident_bytes + need > self.budget:
            victim = next((k for k in self.resident if not protected(k)), None)
            if victim is None:
                return False
            self._drop(victim)
        return True

    def _free(self, entry):
        h, kind, _n, _w, _a, nbytes = entry
        (self.lib.rlg_lu_free if kind == KIND_LU else self.lib.rlg_chol_free)(h)
        self.resident_bytes -= nbytes

    def _drop(self, fid):
        self._free(self.resident.pop(fid))

    def evict(self, fid):
        if fid in self.resident:
            self._drop(fid)

    def close(self):
        """Frees the resident factors and stops the worker threads (the loop
        cannot run afterwards)."""
        for fid in list(self.resident):
            self._drop(fid)
        for w in [self.batch_worker] + self.slot_workers:
            w.jobs.put(None)
        for w in [self.batch_worker] + self.slot_workers:
            w.join(timeout=60)

    # ---------------------------------------------------------------- jobs
    def _eta(self, worker, n, a, nrhs, b, x):
        e = C.c_double()
        st = self.lib.rlg_backward_error_f64_dev(n, a, n, nrhs, b, n, x, n, C.byref(e), self._sptr(worker))
        return e.value if st == 0 else float("inf")

    def _job_batch(self, worker, reqs):
        torch, lib = self.torch, self.lib
        count = len(reqs)
        items = (_lib.BatchItem * count)()
        sp = self._sptr(worker)
        with torch.cuda.stream(worker.stream):
            # plain: the kernel factors in place, so it works on a copy of A;
            # the solution overwrites the copy of B in both forms
            bcat = torch.cat([r.b for r in reqs])
            acat = torch.cat([r.a for r in reqs]) if not self.small_mixed else None
            piv = torch.empty(sum(r.n for r in reqs), dtype=torch.int64, device=self.dev)
            boff = aoff = poff = 0
            for i, r in enumerate(reqs):
                pa = r.a.data_ptr() if self.small_mixed else acat.data_ptr() + 8 * aoff
                items[i] = _lib.BatchItem(r.kind, r.n, r.nrhs, -1, pa, bcat.data_ptr() + 8 * boff,
                                          piv.data_ptr() + 8 * poff)
                aoff += r.n * r.n
                boff += r.nrhs * r.n
                poff += r.n
            host = torch.frombuffer(bytearray(C.string_at(C.addressof(items), C.sizeof(items))),
                                    dtype=torch.uint8)
            dev_items = host.to(self.dev)
            fb = C.c_size_t(0)
            if self.small_mixed:
                st = lib.rlg_batch_solve_mixed_dev(count, dev_items.data_ptr(), 30, self.tol, C.byref(fb), sp)
            else:
                st = lib.rlg_batch_solve_f64_dev(count, dev_items.data_ptr(), sp)
            if st != 0:
                raise RuntimeError(_lib.last_error())
            boff = 0
            for r in reqs:
                r.x.copy_(bcat[boff:boff + r.nrhs * r.n])
                boff += r.nrhs * r.n
            back = dev_items.cpu()  # waits for this worker's stream only
        t = time.perf_counter()
        C.memmove(C.addressof(items), back.numpy().ctypes.data, C.sizeof(items))
        for i, r in enumerate(reqs):
            r.status = items[i].status
            r.done = t
            r.path = "batched_mixed" if self.small_mixed else "batched"
        for r in reqs:
            if r.check and r.status == 0:
                r.eta = self._eta(worker, r.n, r.a.data_ptr(), r.nrhs, r.b.data_ptr(), r.x.data_ptr())

    def _job_single(self, worker, job):
        r, resident = job  # resident: (handle, kind, a) for a reuse request, else None
        lib, torch = self.lib, self.torch
        n = r.n
        sp = self._sptr(worker)
        if resident is not None:
            h, kind, a = resident
            solve = lib.rlg_lu_solve_f64_dev if kind == KIND_LU else lib.rlg_chol_solve_f64_dev
            r.status = solve(h, r.nrhs, r.b.data_ptr(), n, r.x.data_ptr(), n, sp)
            worker.stream.synchronize()
            r.done, r.path = time.perf_counter(), "reuse"
            if r.check and r.status == 0:
                r.eta = self._eta(worker, n, a.data_ptr(), r.nrhs, r.b.data_ptr(), r.x.data_ptr())
            return
        if self.mixed and r.keep_id is None and r.kind == KIND_LU:
            opts = _lib.MixedOpts(_lib.PREC_F16X3, 0, self.tol, 30, 1, self.mixed_pivot)
            rep = _lib.MixedReport()
            r.status = lib.rlg_solve_mixed_f64_dev(n, r.a.data_ptr(), n, r.nrhs, r.b.data_ptr(), n,
                                                   r.x.data_ptr(), n, C.byref(opts), C.byref(rep), sp)
            worker.stream.synchronize()
            r.done, r.path, r.eta, r.fell_back = time.perf_counter(), "single_mixed", rep.backward_error, rep.fell_back
            return
        with torch.cuda.stream(worker.stream):
            work = torch.empty(n * n, dtype=torch.float64, device=self.dev)
        lib.rlg_dev_copy(work.data_ptr(), r.a.data_ptr(), 8 * n * n, sp)
        h = C.c_void_p()
        if r.kind == KIND_LU:
            r.status = lib.rlg_lu_decompose_f64_dev(n, work.data_ptr(), n, sp, C.byref(h))
            if r.status == 0:
                r.status = lib.rlg_lu_solve_f64_dev(h, r.nrhs, r.b.data_ptr(), n, r.x.data_ptr(), n, sp)
        else:
            r.status = lib.rlg_chol_decompose_f64_dev(n, work.data_ptr(), n, sp, C.byref(h))
            if r.status == 0:
                r.status = lib.rlg_chol_solve_f64_dev(h, r.nrhs, r.b.data_ptr(), n, r.x.data_ptr(), n, sp)
        worker.stream.synchronize()
        r.done, r.path = time.perf_counter(), "single"
        if r.check and r.status == 0:
            r.eta = self._eta(worker, n, r.a.data_ptr(), r.nrhs, r.b.data_ptr(), r.x.data_ptr())
        free = lib.rlg_lu_free if r.kind == KIND_LU else lib.rlg_chol_free
        if r.keep_id is not None and r.status == 0 and h.value:
            r._kept = (h, work)  # handed to the dispatcher, which makes it resident
        elif h.value:
            free(h)

    # ---------------------------------------------------------------- loop
    def run(self, requests, clock_start=None):
        """Serve `requests` (their `arrival` is in seconds from the start).
        Returns the requests with status, done, latency and eta filled."""
        t0 = time.perf_counter() if clock_start is None else clock_start
        for r in requests:
            r.t_arr = t0 + r.arrival
        pending = collections.deque(sorted(requests, key=lambda r: r.t_arr))
        small_q = []          # admitted one-shot systems of order <= 512
        single_q = collections.deque()  # admitted per-system factorizations
        reuse_q = collections.deque()   # admitted solves on a resident factor: served first
        queued_reuse = collections.Counter()  # resident id -> queued requests that will use it
        waiting_reuse = []    # reuse requests whose factor is still being computed
        keeping = {}          # keep_id -> in-flight request
        in_flight_bytes = 0
        in_use = collections.Counter()  # resident id -> requests using it
        retired = []          # (id, entry): factors replaced under their id while still in use
        outstanding = len(requests)
        failure = None

        def finish_now(r, status, path):
            nonlocal outstanding
            r.status, r.done, r.path = status, time.perf_counter(), path
            outstanding -= 1

        while outstanding:
            # ---- completions (non-blocking)
            while True:
                try:
                    worker, arg, err = self.completions.get_nowait()
                except queue.Empty:
                    break
                worker.busy = False
                if err is not None and failure is None:
                    failure = err
                done = arg if isinstance(arg, list) else [arg[0]]
                for r in done:
                    outstanding -= 1
                    in_flight_bytes -= r.need
                    if r.reuse_id is not None:
                        in_use[r.reuse_id] -= 1
                    if r.keep_id is not None:
                        keeping.pop(r.keep_id, None)
                        kept, r._kept = r._kept, None
                        if kept is not None:
                            if r.keep_id in self.resident:
                                if in_use[r.keep_id] > 0:  # freed once its solves are done
                                    retired.append((r.keep_id, self.resident.pop(r.keep_id)))
                                else:
                                    self._drop(r.keep_id)
                            nbytes = 16 * r.n * r.n  # the factor and the original matrix
                            self.resident[r.keep_id] = (kept[0], r.kind, r.n, kept[1], r.a, nbytes)
                            self.resident_bytes += nbytes
            if retired:
                keep = []
                for fid, entry in retired:
                    if in_use[fid] > 0:
                        keep.append((fid, entry))
                    else:
                        self._free(entry)
                retired = keep
            if failure is not None:
                raise failure
            # ---- reuse requests whose factor has since been computed
            still = []
            for r in waiting_reuse:
                if r.reuse_id in keeping:
                    still.append(r)
                elif r.reuse_id in self.resident:
                    reuse_q.append(r)
                else:
                    queued_reuse[r.reuse_id] -= 1
                    in_flight_bytes -= r.need
                    finish_now(r, 1, "reuse")  # InvalidArg: the factor is not resident
            waiting_reuse = still
            # ---- admission in arrival order
            now = time.perf_counter()
            while pending and pending[0].t_arr <= now:
                r = pending[0]
                need = self._need(r)
                if need > self.budget:
                    pending.popleft()
                    finish_now(r, RLG_E_OOM, "rejected")
                    continue
                def protected(k, rid=r.reuse_id):
                    return in_use[k] > 0 or queued_reuse[k] > 0 or k == rid
                if in_flight_bytes + need > self.budget or not self._evict_for(in_flight_bytes + need,
                                                                                 protected):
                    break  # wait for work in flight to finish
                pending.popleft()
                r.need = need
                in_flight_bytes += need
                if r.reuse_id is not None:
                    if r.reuse_id in keeping:
                        waiting_reuse.append(r)
                        queued_reuse[r.reuse_id] += 1
                    elif r.reuse_id in self.resident:
                        reuse_q.append(r)
                        queued_reuse[r.reuse_id] += 1
                    else:
                        in_flight_bytes -= need
                        finish_now(r, 1, "reuse")
                elif r.n <= SMALL_MAX and r.keep_id is None:
                    small_q.append(r)
                else:
                    if r.keep_id is not None:
                        keeping[r.keep_id] = r
                    single_q.append(r)
            # ---- dispatch
            if small_q and not self.batch_worker.busy:
                self.batch_worker.busy = True
                self.batch_worker.jobs.put((ServingLoop._job_batch.__get__(self), small_q))
                small_q = []
            for w in self.slot_workers:
                if not reuse_q and not single_q:
                    break
                if w.busy:
                    continue
                r = reuse_q.popleft() if reuse_q else single_q.popleft()
                resident = None
                if r.reuse_id is not None:
                    self.resident.move_to_end(r.reuse_id)
                    h, kind, _rn, _w, a, _nb = self.resident[r.reuse_id]
                    resident = (h, kind, a)
                    in_use[r.reuse_id] += 1
                    queued_reuse[r.reuse_id] -= 1
                w.busy = True
                w.jobs.put((ServingLoop._job_single.__get__(self), (r, resident)))
            # ---- wait for the next event: a completion or the next arrival
            if outstanding:
                timeout = 1e-3
                if pending:
                    timeout = max(0.0, min(timeout, pending[0].t_arr - time.perf_counter()))
                try:
                    item = self.completions.get(timeout=timeout) if timeout > 0 else None
                except queue.Empty:
                    item = None
                if item is not None:
                    self.completions.put(item)  # handled at the top of the loop
        for _fid, entry in retired:  # every request is done: nothing uses them
            self._free(entry)
        return requests


def summarize(reqs):
    """solves/s over the whole run, p50/p99 latency, backward error."""
    import numpy as np
    ok = [r for r in reqs if r.status == 0]
    lat = np.array([r.latency_ms for r in ok]) if ok else np.array([0.0])
    first = min(r.t_arr for r in reqs)
    last = max(r.done for r in reqs if r.done is not None)
    checked = [r for r in ok if r.eta is not None]
    worst = max(checked, key=lambda r: r.eta) if checked else None
    paths = collections.Counter(r.path for r in reqs)
    return {"requests": len(reqs), "completed": len(ok), "failed": len(reqs) - len(ok),
            "solves_per_s": round(len(ok) / (last - first), 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "backward_error_max": worst.eta if worst else None, "eta_checked": len(checked),
            "worst": ({"rid": worst.rid, "kind": "lu" if worst.kind == KIND_LU else "chol", "n": worst.n,
                       "path": worst.path} if worst else None),
            "fallbacks": sum(r.fell_back for r in ok), "paths": dict(paths),
            "by_path": _by_path(ok)}


def _by_path(ok):
    """Per path: requests, and p50 / p99 of the time queued before a worker
    took the request and of the time from then to its result (ms)."""
    import numpy as np
    out = {}
    for path in sorted({r.path for r in ok}):
        rs = [r for r in ok if r.path == path and r.t_start is not None]
        if not rs:
            continue
        q = np.array([(r.t_start - r.t_arr) * 1e3 for r in rs])
        sv = np.array([(r.done - r.t_start) * 1e3 for r in rs])
        out[path] = {"requests": len(rs), "queue_p50_ms": round(float(np.percentile(q, 50)), 2),
                     "queue_p99_ms": round(float(np.percentile(q, 99)), 2),
                     "service_p50_ms": round(float(np.percentile(sv, 50)), 2),
                     "service_p99_ms": round(float(np.percentile(sv, 99)), 2)}
    return out
