"""Host helpers for the batched and serving entry points of librlgpu.

solve_batch(): packs many (kind, A, B) systems into device buffers, runs the
persistent batched kernel (rlg_batch_solve_f64_dev / _mixed_dev) and returns
the solutions, factors and statuses.

Server: a thin wrapper of the continuous-batching request loop
(rlg_server_*).
"""
import ctypes as C

import numpy as np

from . import _lib

KIND_LU, KIND_CHOL = _lib.KIND_LU, _lib.KIND_CHOL


class DeviceArena:
    """One device allocation carved into 256-byte-aligned pieces."""

    def __init__(self, nbytes):
        self.lib = _lib.load()
        self.base = C.c_void_p()
        st = self.lib.rlg_dev_alloc(max(nbytes, 256), C.byref(self.base))
        if st != 0:
            raise MemoryError(_lib.last_error())
        self.off = 0
        self.size = nbytes

    def take(self, nbytes):
        p = self.base.value + self.off
        self.off += (nbytes + 255) & ~255
        assert self.off <= self.size
        return p

    def free(self):
        if self.base:
            self.lib.rlg_dev_free(self.base)
            self.base = C.c_void_p()


def solve_batch(systems, mixed=False, max_iter=30, tol=1e-13):
    """systems: list of (kind, A (n x n), B (nrhs x n)).  Returns dict with
    'x' (list of (nrhs, n) arrays), 'factor' (packed factors; for mixed the
    input is kept), 'ipiv', 'status', 'fallbacks'."""
    lib = _lib.load()
    sizes = []
    total = 0
    for kind, a, b in systems:
        n = a.shape[0]
        nr = b.shape[0]
        sizes.append((n, nr))
        total += ((n * n * 8 + 255) & ~255) + ((nr * n * 8 + 255) & ~255) + ((n * 8 + 255) & ~255)
    count = len(systems)
    total += count * C.sizeof(_lib.BatchItem) + 256
    arena = DeviceArena(total)
    items = (_lib.BatchItem * count)()
    ptrs = []
    try:
        for i, (kind, a, b) in enumerate(systems):
            n, nr = sizes[i]
            a = np.ascontiguousarray(a, dtype=np.float64)
            b = np.ascontiguousarray(b, dtype=np.float64)
            pa, pb, pp = arena.take(n * n * 8), arena.take(nr * n * 8), arena.take(n * 8)
            assert lib.rlg_dev_upload(pa, a.ctypes.data, a.nbytes) == 0
            assert lib.rlg_dev_upload(pb, b.ctypes.data, b.nbytes) == 0
            items[i] = _lib.BatchItem(kind, n, nr, -1, pa, pb, pp)
            ptrs.append((pa, pb, pp))
        pitems = arena.take(C.sizeof(items))
        assert lib.rlg_dev_upload(pitems, C.addressof(items), C.sizeof(items)) == 0
        fb = C.c_size_t(0)
        if mixed:
            st = lib.rlg_batch_solve_mixed_dev(count, pitems, max_iter, tol, C.byref(fb), None)
        else:
            st = lib.rlg_batch_solve_f64_dev(count, pitems, None)
        if st != 0:
            raise RuntimeError(_lib.last_error())
        assert lib.rlg_synchronize() == 0
        assert lib.rlg_dev_download(C.addressof(items), pitems, C.sizeof(items)) == 0
        out = {"x": [], "factor": [], "ipiv": [], "status": [], "fallbacks": fb.value}
        for i, (pa, pb, pp) in enumerate(ptrs):
            n, nr = sizes[i]
            x = np.empty((nr, n))
            f = np.empty((n, n))
            ip = np.empty(n, dtype=np.int64)
            assert lib.rlg_dev_download(x.ctypes.data, pb, x.nbytes) == 0
            assert lib.rlg_dev_download(f.ctypes.data, pa, f.nbytes) == 0
            assert lib.rlg_dev_download(ip.ctypes.data, pp, ip.nbytes) == 0
            out["x"].append(x)
            out["factor"].append(f)
            out["ipiv"].append(ip)
            out["status"].append(items[i].status)
        return out
    finally:
        arena.free()


class Server:
    """Continuous-batching solve server (rlg_server_*)."""

    def __init__(self, mem_budget_bytes=64 << 30, slots=4, mixed=False, tol=1e-13):
        self.lib = _lib.load()
        o = _lib.ServerOpts(mem_budget_bytes, slots, 1 if mixed else 0, tol)
        self.h = C.c_void_p()
        st = self.lib.rlg_server_create(C.byref(o), C.byref(self.h))
        if st != 0:
            raise RuntimeError(_lib.last_error())
        self._keep = {}

    def submit(self, kind, a, b, factor_id=0):
        """a: (n, n) or None to reuse factor_id; b: (nrhs, n).  Returns a ticket."""
        b = np.ascontiguousarray(np.atleast_2d(b), dtype=np.float64)
        n = b.shape[1]
        pa = None
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.float64)
            pa = a.ctypes.data_as(_lib.PD)
        req = _lib.Request(kind, n, b.shape[0], pa, b.ctypes.data_as(_lib.PD), factor_id)
        t = C.c_int64()
        st = self.lib.rlg_server_submit(self.h, C.byref(req), C.byref(t))
        if st != 0:
            raise RuntimeError(_lib.last_error())
        self._keep[t.value] = (a, b)  # buffers must outlive the request
        return t.value

    def wait(self, ticket):
        r = _lib.Result()
        st = self.lib.rlg_server_wait(self.h, ticket, C.byref(r))
        if st != 0:
            raise RuntimeError(_lib.last_error())
        a, b = self._keep.pop(ticket)
        x = None
        if r.x:
            x = np.ctypeslib.as_array(r.x, shape=(b.shape[0] * b.shape[1],)).copy().reshape(b.shape)
        self.lib.rlg_server_release(self.h, ticket)
        return {"status": r.status, "x": x, "backward_error": r.backward_error,
                "latency_ms": r.latency_ms, "fell_back": r.fell_back}

    def evict(self, factor_id):
        return self.lib.rlg_server_evict(self.h, factor_id)

    def close(self):
        if self.h:
            self.lib.rlg_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()
