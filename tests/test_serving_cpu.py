"""CPU: the serving loop's host logic (rulinalg_amd.serving) with a stub
library: routing (batched launch for one-shot orders <= 512, per-system path
otherwise), admission by the memory budget (RLG_E_OOM for a request that can
never fit), least-recently-used eviction of resident factors, and reuse of a
resident factor.  Device pointers are host pointers here; the stub records
calls instead of computing."""
import contextlib
import ctypes as C
import types

import pytest

torch = pytest.importorskip("torch")

from rulinalg_amd import _lib, serving  # noqa: E402


class StubLib:
    def __init__(self):
        self.calls = []
        self.handles = {}
        self.next = 1

    def _h(self, out, kind, n):
        h = self.next
        self.next += 1
        self.handles[h] = (kind, n)
        out._obj.value = h
        return 0

    # memory
    def rlg_dev_upload(self, dst, src, n):
        C.memmove(dst, src, n)
        return 0

    def rlg_dev_download(self, dst, src, n):
        C.memmove(dst, src, n)
        return 0

    def rlg_synchronize(self):
        return 0

    def rlg_dev_copy(self, dst, src, n, stream):
        C.memmove(dst, src, n)
        return 0

    # solvers
    def rlg_batch_solve_f64_dev(self, count, items, stream):
        arr = (_lib.BatchItem * count).from_address(items)
        self.calls.append(("batch", [arr[i].n for i in range(count)]))
        for i in range(count):
            arr[i].status = 0
        return 0

    def rlg_batch_solve_mixed_dev(self, count, items, max_iter, tol, fb, stream):
        return self.rlg_batch_solve_f64_dev(count, items, stream)

    def rlg_lu_decompose_f64_dev(self, n, a, lda, stream, out):
        self.calls.append(("lu", n))
        return self._h(out, "lu", n)

    def rlg_chol_decompose_f64_dev(self, n, a, lda, stream, out):
        self.calls.append(("chol", n))
        return self._h(out, "chol", n)

    def rlg_lu_solve_f64_dev(self, h, nrhs, b, ldb, x, ldx, stream):
        self.calls.append(("lu_solve", self.handles[h.value if hasattr(h, "value") else h][1]))
        return 0

    rlg_chol_solve_f64_dev = rlg_lu_solve_f64_dev

    def rlg_lu_free(self, h):
        self.calls.append(("free", h.value if hasattr(h, "value") else h))

    rlg_chol_free = rlg_lu_free

    def rlg_backward_error_f64_dev(self, n, a, lda, nrhs, b, ldb, x, ldx, eta, stream):
        eta._obj.value = 1e-16
        return 0

    def rlg_solve_mixed_f64_dev(self, n, a, lda, nrhs, b, ldb, x, ldx, opts, rep, stream):
        self.calls.append(("mixed", n))
        rep._obj.backward_error = 1e-15
        return 0


class HostStream:
    """Stands in for torch.cuda.Stream: work on the host is always complete."""
    cuda_stream = 0

    def __init__(self, device=None):
        pass

    def synchronize(self):
        pass


@contextlib.contextmanager
def host_stream_ctx(stream):
    yield stream


def torch_cpu():
    """torch with the device calls the loop makes mapped onto the CPU."""
    t = types.SimpleNamespace()
    t.cat = torch.cat
    t.frombuffer = torch.frombuffer
    t.empty = lambda *a, **k: torch.empty(*a, **{kk: v for kk, v in k.items() if kk != "device"})
    t.uint8, t.int64, t.float64 = torch.uint8, torch.int64, torch.float64
    t.cuda = types.SimpleNamespace(synchronize=lambda: None, Stream=HostStream, stream=host_stream_ctx)
    return t


def req(rid, kind, n, arrival=0.0, **kw):
    a = torch.zeros(n * n, dtype=torch.float64)
    b = torch.ones(n, dtype=torch.float64)
    return serving.Request(rid, kind, n, 1, a, b, torch.empty_like(b), arrival, **kw)


def test_routing_batch_and_single():
    lib = StubLib()
    loop = serving.ServingLoop(lib, torch_cpu(), "cpu")
    reqs = [req(0, serving.KIND_LU, 100), req(1, serving.KIND_CHOL, 512), req(2, serving.KIND_LU, 513),
            req(3, serving.KIND_CHOL, 900)]
    loop.run(reqs)
    assert [r.path for r in reqs] == ["batched", "batched", "single", "single"]
    assert all(r.status == 0 and r.eta is not None for r in reqs)
    assert ("batch", [100, 512]) in lib.calls
    assert ("lu", 513) in lib.calls and ("chol", 900) in lib.calls
    # one-shot per-system handles are freed
    assert sum(1 for c in lib.calls if c[0] == "free") == 2
    s = serving.summarize(reqs)
    assert s["completed"] == 4 and s["paths"] == {"batched": 2, "single": 2}


def test_budget_rejects_what_never_fits():
    lib = StubLib()
    loop = serving.ServingLoop(lib, torch_cpu(), "cpu", mem_budget_bytes=1 << 20)
    big, small = req(0, serving.KIND_LU, 400), req(1, serving.KIND_LU, 100, arrival=0.0)
    loop.run([big, small])
    assert big.status == serving.RLG_E_OOM and big.path == "rejected"
    assert small.status == 0


def test_resident_reuse_and_lru_eviction():
    lib = StubLib()
    # room for two resident 128-order factors (16 n^2 bytes each) plus a pass
    budget = 2 * 16 * 128 * 128 + 8 * 128 * 129
    loop = serving.ServingLoop(lib, torch_cpu(), "cpu", mem_budget_bytes=budget)
    k1 = req(0, serving.KIND_LU, 128, keep_id=1)
    k2 = req(1, serving.KIND_LU, 128, arrival=0.001, keep_id=2)
    loop.run([k1, k2])
    assert set(loop.resident) == {1, 2}
    use1 = req(2, serving.KIND_LU, 128, reuse_id=1)
    loop.run([use1])  # factor 1 becomes the most recently used
    assert use1.path == "reuse" and use1.status == 0
    k3 = req(3, serving.KIND_LU, 128, keep_id=3)
    loop.run([k3])  # needs room: factor 2, the least recently used, goes
    assert set(loop.resident) == {1, 3}
    late = req(4, serving.KIND_LU, 128, reuse_id=2)
    loop.run([late])
    assert late.status == 1  # InvalidArg: no longer resident
    loop.close()
    assert not loop.resident and loop.resident_bytes == 0


def test_mixed_mode_routes_large_lu_to_the_mixed_solve():
    """Mixed mode: one-shot LU above order 512 takes the mixed solve,
    Cholesky the FP64 path, and small systems the FP64 batched launch unless
    `small_mixed` asks for the mixed batched one."""
    for small_mixed, small_path in ((False, "batched"), (True, "batched_mixed")):
        lib = StubLib()
        loop = serving.ServingLoop(lib, torch_cpu(), "cpu", mixed=True, small_mixed=small_mixed)
        reqs = [req(0, serving.KIND_LU, 600), req(1, serving.KIND_CHOL, 600), req(2, serving.KIND_LU, 64)]
        loop.run(reqs)
        assert [r.path for r in reqs] == ["single_mixed", "single", small_path]
        assert ("mixed", 600) in lib.calls
        loop.close()


def test_replaced_factor_outlives_its_solves():
    """A request that keeps its factor under an id that is already resident
    replaces it; a reuse solve still running on the old factor keeps it: the
    old handle is freed only after that solve has finished."""
    import threading as th

    gate = th.Event()

    class GatedLib(StubLib):
        def rlg_lu_solve_f64_dev(self, h, nrhs, b, ldb, x, ldx, stream):
            hv = h.value if hasattr(h, "value") else h
            # the reuse request's solve: the second one on the first factor
            if hv == 1 and ("solve_1",) in self.calls and not gate.is_set():
                self.calls.append(("solve_start", hv))
                gate.wait(5)
                self.calls.append(("solve_done", hv))
                return 0
            if hv == 1:
                self.calls.append(("solve_1",))
            return super().rlg_lu_solve_f64_dev(h, nrhs, b, ldb, x, ldx, stream)

    lib = GatedLib()
    loop = serving.ServingLoop(lib, torch_cpu(), "cpu")
    first = req(0, serving.KIND_LU, 600, keep_id=7)
    reuse = req(1, serving.KIND_LU, 600, arrival=0.05, reuse_id=7)
    second = req(2, serving.KIND_LU, 600, arrival=0.10, keep_id=7)
    timer = th.Timer(0.6, gate.set)
    timer.start()
    loop.run([first, reuse, second])
    timer.join()
    assert all(r.status == 0 for r in (first, reuse, second))
    assert ("solve_done", 1) in lib.calls
    # the first factor (handle 1) is freed, and only after the solve on it ended
    assert lib.calls.index(("free", 1)) > lib.calls.index(("solve_done", 1))
    assert 7 in loop.resident and loop.resident[7][0].value == 2
    loop.close()


def test_reuse_solves_go_before_queued_factorizations():
    """With one slot worker busy, a reuse solve admitted after a large
    factorization is still dispatched before it."""
    import time as _t

    class SlowLib(StubLib):
        def rlg_lu_decompose_f64_dev(self, n, a, lda, stream, out):
            if n == 600:
                _t.sleep(0.3)  # the slot stays busy while the next two arrive
            return super().rlg_lu_decompose_f64_dev(n, a, lda, stream, out)

    lib = SlowLib()
    loop = serving.ServingLoop(lib, torch_cpu(), "cpu", slots=1)
    keep = req(0, serving.KIND_LU, 600, keep_id=5)
    big = req(1, serving.KIND_LU, 700, arrival=0.05)
    again = req(2, serving.KIND_LU, 600, arrival=0.1, reuse_id=5)
    loop.run([keep, big, again])
    assert all(r.status == 0 for r in (keep, big, again))
    assert again.path == "reuse"
    solves = [i for i, c in enumerate(lib.calls) if c == ("lu_solve", 600)]
    assert len(solves) == 2  # the keep request's own solve, then the reuse solve
    assert solves[1] < lib.calls.index(("lu", 700))
    loop.close()
