"""GPU: factorizations issued from several host threads at once, each thread
on its own stream, give the same bits as the same factorizations run one at
a time on the library's stream.

The reference's decompositions are plain values, so any number of them may
be computed side by side (`lu.rs:163-244`, `cholesky.rs:116-202`); a
drop-in must keep every result independent of what else runs on the device.
LU's panels are cooperative launches, Cholesky's kernels are ordinary ones;
both are covered, at orders whose panels span many workgroups.
"""
import ctypes as C
import os
import threading
import time

import numpy as np
import pytest

from rulinalg_amd import _lib, workloads as wl

pytestmark = pytest.mark.gpu

KIND_LU, KIND_CHOL = _lib.KIND_LU, _lib.KIND_CHOL


@pytest.fixture(scope="module")
def env():
    import torch
    lib = _lib.load()
    assert lib.rlg_set_device(0) == 0, _lib.last_error()
    return lib, torch, torch.device("cuda:0")


HELPERS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_helpers", "libocc.so")


def factor(lib, torch, kind, a, n, stream):
    """Copy of `a` factored in place on `stream` (None: the library's own);
    returns (factor tensor, pivot vector or None, status)."""
    work = torch.empty_like(a)
    sp = C.c_void_p(stream.cuda_stream) if stream is not None else None
    assert lib.rlg_dev_copy(work.data_ptr(), a.data_ptr(), 8 * n * n, sp) == 0
    h = C.c_void_p()
    dec = lib.rlg_lu_decompose_f64_dev if kind == KIND_LU else lib.rlg_chol_decompose_f64_dev
    st = dec(n, work.data_ptr(), n, sp, C.byref(h))
    piv = None
    if st == 0 and kind == KIND_LU:
        piv = np.empty(n, dtype=np.int64)
        assert lib.rlg_lu_ipiv(h, piv.ctypes.data) == 0
    if stream is not None:
        stream.synchronize()
    else:
        assert lib.rlg_synchronize() == 0
    if h.value:
        (lib.rlg_lu_free if kind == KIND_LU else lib.rlg_chol_free)(h)
    return work, piv, st


def test_factorizations_on_concurrent_streams_match_alone(env):
    lib, torch, dev = env
    r = wl.rng(97)
    # six threads: four LU (cooperative panels of 8 to 24 workgroups), two Cholesky
    plan = [(KIND_LU, 1024), (KIND_LU, 2048), (KIND_LU, 3072), (KIND_LU, 1536),
            (KIND_CHOL, 2048), (KIND_CHOL, 2560)]
    mats = []
    for kind, n in plan:
        g = r.standard_normal((n, n))
        a = g if kind == KIND_LU else g @ g.T / n + np.eye(n)
        mats.append(torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).to(dev))
    torch.cuda.synchronize()
    alone = []
    for (kind, n), a in zip(plan, mats):
        work, piv, st = factor(lib, torch, kind, a, n, None)
        assert st == 0, _lib.last_error()
        alone.append((work, piv))

    rounds = 4
    results = [[] for _ in plan]
    errors = []
    start = threading.Barrier(len(plan))

    def run(i):
        try:
            kind, n = plan[i]
            s = torch.cuda.Stream(device=dev)
            start.wait()
            for _ in range(rounds):
                with torch.cuda.stream(s):
                    work, piv, st = factor(lib, torch, kind, mats[i], n, s)
                    same = st == 0 and bool(torch.equal(work, alone[i][0]))
                    diff = 0.0 if same or st else float((work - alone[i][0]).abs().max())
                s.synchronize()
                same_piv = piv is None or np.array_equal(piv, alone[i][1])
                results[i].append((st, same, same_piv, diff))
        except Exception as e:  # reported below, with the thread's case
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(plan))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads), "a factorization thread did not finish"
    assert not errors, errors
    bad = [(("lu" if plan[i][0] == KIND_LU else "chol"), plan[i][1], k, res)
           for i in range(len(plan)) for k, res in enumerate(results[i])
           if res[0] != 0 or not res[1] or not res[2]]
    assert not bad, f"(kind, n, round, (status, same factor, same pivots, max |diff|)): {bad}"


def test_cholesky_block_columns_starting_in_waves(env):
    """Cholesky of order 2048 while a kernel on another stream holds all but
    one CU of each XCD: each 64-column block launch (31 workgroups for the
    first) then runs a few workgroups at a time, so its first workgroups end
    before its last ones start.  Every workgroup must still read the
    unfactored diagonal block, which is written back in place once all of
    them have read it: the factor must equal, bit for bit, the one computed
    with the whole device free."""
    lib, torch, dev = env
    occ = C.CDLL(HELPERS)
    occ.occ_launch.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p]
    occ.occ_stream_create.argtypes = [C.POINTER(C.c_void_p)]
    occ.occ_stream_destroy.argtypes = [C.c_void_p]
    occ.occ_stream_sync.argtypes = [C.c_void_p]
    r = wl.rng(98)
    n = 2048
    g = r.standard_normal((n, n))
    a = torch.from_numpy(np.ascontiguousarray(g @ g.T / n + np.eye(n)).reshape(-1)).to(dev)
    torch.cuda.synchronize()
    ref, _, st = factor(lib, torch, KIND_CHOL, a, n, None)
    assert st == 0, _lib.last_error()
    t0 = time.perf_counter()
    ref2, _, _ = factor(lib, torch, KIND_CHOL, a, n, None)
    alone_ms = (time.perf_counter() - t0) * 1e3
    held = 256 - 8  # one CU of each of the 8 XCDs stays free (workgroups go round-robin over XCDs)
    sink = torch.zeros(held, dtype=torch.float32, device=dev)
    hog = C.c_void_p()
    assert occ.occ_stream_create(C.byref(hog)) == 0
    torch.cuda.synchronize()
    hold_ms = 100.0
    try:
        assert occ.occ_launch(hog, held, hold_ms, sink.data_ptr()) == 0
        time.sleep(0.005)  # the holding workgroups are resident
        work = torch.empty_like(a)
        assert lib.rlg_dev_copy(work.data_ptr(), a.data_ptr(), 8 * n * n, None) == 0
        h = C.c_void_p()
        t0 = time.perf_counter()
        st = lib.rlg_chol_decompose_f64_dev(n, work.data_ptr(), n, None, C.byref(h))  # returns when done
        held_ms = (time.perf_counter() - t0) * 1e3
        if h.value:
            lib.rlg_chol_free(h)
    finally:
        assert occ.occ_stream_sync(hog) == 0
        occ.occ_stream_destroy(hog)
    assert lib.rlg_synchronize() == 0
    assert st == 0, _lib.last_error()
    lo = torch.tril(work.view(n, n))
    same = bool(torch.equal(lo, torch.tril(ref.view(n, n))))
    diff = float((lo - torch.tril(ref.view(n, n))).abs().max())
    print(f"cholesky {n}: {alone_ms:.2f} ms with the device free, {held_ms:.2f} ms beside the "
          f"holding kernel ({hold_ms:.0f} ms); max |diff| {diff:.3e}")
    assert torch.equal(torch.tril(ref2.view(n, n)), torch.tril(ref.view(n, n)))
    assert same, f"factor differs when its workgroups start in waves: max |diff| {diff:.3e}"
