"""Continuous-batching request loop over the batched and per-system solvers.

One host thread drives the library's stream.  Every pass of the loop admits,
in arrival order, each request that has arrived by now and whose device bytes
fit under the memory budget together with the resident factors (the
least-recently-used resident factor that no admitted request needs is evicted
to make room; a request that can never fit fails with RLG_E_OOM).  The
admitted requests then run:

* one-shot systems of order <= 512 together in ONE persistent batched launch
  (``rlg_batch_solve_f64_dev``, or ``rlg_batch_solve_mixed_dev`` in mixed
  mode): their work overlaps on the device;
* larger systems, and requests that keep or reuse a resident factor, one at a
  time through the per-system path (``rlg_{lu,chol}_decompose_f64_dev`` +
  ``_solve_f64_dev``, or ``rlg_solve_mixed_f64_dev`` for one-shot mixed
  requests).

Requests that arrive while a pass runs are admitted by the next pass.  A
request that keeps its factor (``keep_id``) leaves the handle resident under
that id, with its original matrix for the backward error; a later request with
``reuse_id`` solves new right-hand sides against it without refactoring.

The reference has no serving layer (rulinalg is a library of synchronous
calls); the loop serves its calls: PartialPivLu::decompose + solve
(src/matrix/decomposition/lu.rs:163-244) and Cholesky::decompose + solve
(src/matrix/decomposition/cholesky.rs:116-202).  Results do not depend on what
else is in the pass: the batched kernel solves every system in one workgroup
with a fixed operation order, and the per-system path is deterministic.
"""
import collections
import ctypes as C
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
                 "check", "status", "done", "eta", "path", "fell_back", "t_arr")

    def __init__(self, rid, kind, n, nrhs, a, b, x, arrival, keep_id=None, reuse_id=None, check=True):
        self.rid, self.kind, self.n, self.nrhs = rid, kind, n, nrhs
        self.a, self.b, self.x, self.arrival = a, b, x, arrival
        self.keep_id, self.reuse_id, self.check = keep_id, reuse_id, check
        self.status, self.done, self.eta, self.path, self.fell_back = None, None, None, None, 0
        self.t_arr = None  # arrival on the host clock, set by ServingLoop.run

    @property
    def latency_ms(self):
        return None if self.done is None else (self.done - self.t_arr) * 1e3


class ServingLoop:
    def __init__(self, lib, torch, dev, mem_budget_bytes=64 << 30, mixed=False, tol=1e-13):
        self.lib, self.torch, self.dev = lib, torch, dev
        self.budget = mem_budget_bytes
        self.mixed = mixed
        self.tol = tol
        self.resident = collections.OrderedDict()  # id -> (handle, kind, n, work, a, bytes)
        self.resident_bytes = 0

    # ---------------------------------------------------------------- memory
    @staticmethod
    def _need(r):
        """Device bytes a request holds while it runs (the factor copy, or the
        batch's copies of A and B), plus what stays resident if kept."""
        if r.reuse_id is not None:
            return 0
        return 8 * r.n * r.n + 8 * r.nrhs * r.n

    def _evict_for(self, need, protect):
        while self.resident_bytes + need > self.budget:
            victim = next((k for k in self.resident if k not in protect), None)
            if victim is None:
                return False
            self._drop(victim)
        return True

    def _drop(self, fid):
        h, kind, _n, _w, _a, nbytes = self.resident.pop(fid)
        (self.lib.rlg_lu_free if kind == KIND_LU else self.lib.rlg_chol_free)(h)
        self.resident_bytes -= nbytes

    def evict(self, fid):
        if fid in self.resident:
            self._drop(fid)

    def close(self):
        for fid in list(self.resident):
            self._drop(fid)

    # ---------------------------------------------------------------- passes
    def _eta(self, n, a, nrhs, b, x):
        e = C.c_double()
        st = self.lib.rlg_backward_error_f64_dev(n, a, n, nrhs, b, n, x, n, C.byref(e), None)
        return e.value if st == 0 else float("inf")

    def _run_batch(self, reqs):
        torch, lib = self.torch, self.lib
        count = len(reqs)
        items = (_lib.BatchItem * count)()
        # plain: the kernel factors in place, so it works on a copy of A; the
        # solution overwrites the copy of B in both forms
        bcat = torch.cat([r.b for r in reqs])
        acat = torch.cat([r.a for r in reqs]) if not self.mixed else None
        piv = torch.empty(sum(r.n for r in reqs), dtype=torch.int64, device=self.dev)
        boff = aoff = poff = 0
        for i, r in enumerate(reqs):
            pa = r.a.data_ptr() if self.mixed else acat.data_ptr() + 8 * aoff
            items[i] = _lib.BatchItem(r.kind, r.n, r.nrhs, -1, pa, bcat.data_ptr() + 8 * boff,
                                      piv.data_ptr() + 8 * poff)
            aoff += r.n * r.n
            boff += r.nrhs * r.n
            poff += r.n
        dev_items = torch.empty(C.sizeof(items), dtype=torch.uint8, device=self.dev)
        torch.cuda.synchronize()  # torch's copies are on its own stream
        assert lib.rlg_dev_upload(dev_items.data_ptr(), C.addressof(items), C.sizeof(items)) == 0
        fb = C.c_size_t(0)
        if self.mixed:
            st = lib.rlg_batch_solve_mixed_dev(count, dev_items.data_ptr(), 30, self.tol, C.byref(fb), None)
        else:
            st = lib.rlg_batch_solve_f64_dev(count, dev_items.data_ptr(), None)
        if st != 0:
            raise RuntimeError(_lib.last_error())
        assert lib.rlg_synchronize() == 0
        t = time.perf_counter()
        assert lib.rlg_dev_download(C.addressof(items), dev_items.data_ptr(), C.sizeof(items)) == 0
        boff = 0
        for i, r in enumerate(reqs):
            sol = bcat[boff:boff + r.nrhs * r.n]
            boff += r.nrhs * r.n
            r.status = items[i].status
            r.done = t
            r.path = "batched_mixed" if self.mixed else "batched"
            r.x.copy_(sol)
        torch.cuda.synchronize()
        for r in reqs:
            if r.check and r.status == 0:
                r.eta = self._eta(r.n, r.a.data_ptr(), r.nrhs, r.b.data_ptr(), r.x.data_ptr())

    def _run_single(self, r):
        lib, torch = self.lib, self.torch
        n = r.n
        if r.reuse_id is not None:
            if r.reuse_id not in self.resident:
                r.status, r.done, r.path = 1, time.perf_counter(), "reuse"  # InvalidArg: not resident
                return
            self.resident.move_to_end(r.reuse_id)
            h, kind, _rn, _w, a, _nb = self.resident[r.reuse_id]
            solve = lib.rlg_lu_solve_f64_dev if kind == KIND_LU else lib.rlg_chol_solve_f64_dev
            r.status = solve(h, r.nrhs, r.b.data_ptr(), n, r.x.data_ptr(), n, None)
            assert lib.rlg_synchronize() == 0
            r.done, r.path = time.perf_counter(), "reuse"
            if r.check and r.status == 0:
                r.eta = self._eta(n, a.data_ptr(), r.nrhs, r.b.data_ptr(), r.x.data_ptr())
            return
        if self.mixed and r.keep_id is None and r.kind == KIND_LU:
            opts = _lib.MixedOpts(1, 0, self.tol, 30, 1)
            rep = _lib.MixedReport()
            r.status = lib.rlg_solve_mixed_f64_dev(n, r.a.data_ptr(), n, r.nrhs, r.b.data_ptr(), n,
                                                   r.x.data_ptr(), n, C.byref(opts), C.byref(rep), None)
            assert lib.rlg_synchronize() == 0
            r.done, r.path, r.eta, r.fell_back = time.perf_counter(), "single_mixed", rep.backward_error, rep.fell_back
            return
        work = torch.empty(n * n, dtype=torch.float64, device=self.dev)
        work.copy_(r.a)
        torch.cuda.synchronize()
        h = C.c_void_p()
        if r.kind == KIND_LU:
            r.status = lib.rlg_lu_decompose_f64_dev(n, work.data_ptr(), n, None, C.byref(h))
            if r.status == 0:
                r.status = lib.rlg_lu_solve_f64_dev(h, r.nrhs, r.b.data_ptr(), n, r.x.data_ptr(), n, None)
        else:
            r.status = lib.rlg_chol_decompose_f64_dev(n, work.data_ptr(), n, None, C.byref(h))
            if r.status == 0:
                r.status = lib.rlg_chol_solve_f64_dev(h, r.nrhs, r.b.data_ptr(), n, r.x.data_ptr(), n, None)
        assert lib.rlg_synchronize() == 0
        r.done, r.path = time.perf_counter(), "single"
        if r.check and r.status == 0:
            r.eta = self._eta(n, r.a.data_ptr(), r.nrhs, r.b.data_ptr(), r.x.data_ptr())
        free = lib.rlg_lu_free if r.kind == KIND_LU else lib.rlg_chol_free
        if r.keep_id is not None and r.status == 0 and h.value:
            if r.keep_id in self.resident:
                self._drop(r.keep_id)
            nbytes = 16 * n * n  # the factor and the original matrix
            self.resident[r.keep_id] = (h, r.kind, n, work, r.a, nbytes)
            self.resident_bytes += nbytes
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
        while pending:
            now = time.perf_counter()
            if pending[0].t_arr > now:
                time.sleep(min(pending[0].t_arr - now, 1e-3))
                continue
            admitted = []
            used = 0
            while pending and pending[0].t_arr <= now:
                r = pending[0]
                need = self._need(r) + (16 * r.n * r.n if r.keep_id is not None else 0)
                protect = {q.reuse_id for q in admitted if q.reuse_id is not None}
                if r.reuse_id is not None:
                    protect.add(r.reuse_id)
                if need > self.budget:
                    pending.popleft()
                    r.status, r.done, r.path = RLG_E_OOM, time.perf_counter(), "rejected"
                    continue
                if not self._evict_for(used + need, protect):
                    break  # wait for this pass to finish
                pending.popleft()
                admitted.append(r)
                used += need
            small = [r for r in admitted
                     if r.n <= SMALL_MAX and r.keep_id is None and r.reuse_id is None]
            singles = [r for r in admitted if r not in small]
            if small:
                self._run_batch(small)
            for r in singles:
                self._run_single(r)
        return requests


def summarize(reqs):
    """solves/s over the whole run, p50/p99 latency, backward error."""
    import numpy as np
    ok = [r for r in reqs if r.status == 0]
    lat = np.array([r.latency_ms for r in ok]) if ok else np.array([0.0])
    first = min(r.t_arr for r in reqs)
    last = max(r.done for r in reqs if r.done is not None)
    etas = [r.eta for r in ok if r.eta is not None]
    paths = collections.Counter(r.path for r in reqs)
    return {"requests": len(reqs), "completed": len(ok), "failed": len(reqs) - len(ok),
            "solves_per_s": round(len(ok) / (last - first), 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "backward_error_max": max(etas) if etas else None, "eta_checked": len(etas),
            "fallbacks": sum(r.fell_back for r in ok), "paths": dict(paths)}
