"""GPU: the continuous-batching request loop (rulinalg_amd.serving).

Every result must have status 0 and a normwise backward error <= 1e-13
(measured on the device for every request), agree with the CPU oracle within
the single-system tolerance, and be bit-identical whether the request ran
alone or in a mix.  Admission: a request larger than the memory budget fails
with RLG_E_OOM while smaller ones are served; a resident factor serves later
right-hand sides until it is evicted.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from rulinalg_amd import _lib, serving, workloads as wl

pytestmark = pytest.mark.gpu
U = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def env():
    import torch
    lib = _lib.load()
    assert lib.rlg_set_device(0) == 0, _lib.last_error()
    return lib, torch, torch.device("cuda:0")


def system(r, kind, n, nrhs):
    if kind == serving.KIND_LU:
        a = r.standard_normal((n, n))
    else:
        g = r.standard_normal((n, n))
        a = g @ g.T / n + np.eye(n)
    return a, r.standard_normal((nrhs, n))


def request(torch, dev, rid, kind, a, b, arrival=0.0, **kw):
    at = torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).to(dev)
    bt = torch.from_numpy(np.ascontiguousarray(b).reshape(-1)).to(dev)
    return serving.Request(rid, kind, a.shape[0], b.shape[0], at, bt, torch.empty_like(bt), arrival, **kw)


def host_eta(a, x, b):
    return np.abs(b - a @ x).max() / (np.abs(a).sum(1).max() * np.abs(x).max() + np.abs(b).max())


def test_mix_against_oracle_and_alone(env):
    lib, torch, dev = env
    r = wl.rng(81)
    specs = []
    for i in range(24):
        n = int(np.exp(r.uniform(np.log(4), np.log(700))))
        kind = serving.KIND_LU if i % 3 else serving.KIND_CHOL
        a, b = system(r, kind, n, int(r.integers(1, 4)))
        specs.append((kind, a, b, 0.001 * i))
    loop = serving.ServingLoop(lib, torch, dev)
    mix = [request(torch, dev, i, k, a, b, t) for i, (k, a, b, t) in enumerate(specs)]
    loop.run(mix)
    torch.cuda.synchronize()
    paths = {q.path for q in mix}
    assert "batched" in paths and "single" in paths
    for (kind, a, b, _), q in zip(specs, mix):
        assert q.status == 0, (q.rid, q.n, q.path)
        assert q.eta is not None and q.eta <= 1e-13
        x = q.x.cpu().numpy().reshape(b.shape)
        k = np.linalg.cond(a, np.inf)
        for i in range(b.shape[0]):
            assert host_eta(a, x[i], b[i]) <= 1e-13
            if kind == serving.KIND_LU:
                lu, perm, _ = orc.lu_decompose(a)
                xo = orc.lu_solve(lu, perm, b[i])
            else:
                xo = orc.chol_solve(orc.chol_decompose(a), b[i])
            assert np.abs(x[i] - xo).max() / np.abs(xo).max() <= max(1e-13, 16 * a.shape[0] * k * U)
    # each request alone (the only request of a run) gives the same bits
    solo = serving.ServingLoop(lib, torch, dev)
    for (kind, a, b, _), q in zip(specs, mix):
        alone = request(torch, dev, q.rid, kind, a, b)
        solo.run([alone])
        torch.cuda.synchronize()
        assert alone.status == 0 and alone.path == q.path
        assert torch.equal(alone.x, q.x)
    solo.close()
    loop.close()


@pytest.mark.parametrize("mixed", [False, True], ids=["plain", "mixed"])
def test_requests_overlap_and_keep_their_results(env, mixed):
    """Continuous admission: 40 requests arrive over 20 ms while earlier ones
    are on the GPU (orders 600-1500 on four slot workers, small ones batched
    meanwhile); every result equals the one its request gets alone, and the
    run takes less time than the same requests one after another.  In mixed
    mode the large LU systems are factored with tournament pivoting (the
    loop's default), whose result does not depend on which workgroup merges
    a tree node."""
    lib, torch, dev = env
    r = wl.rng(85)
    specs = []
    for i in range(40):
        n = int(r.integers(600, 1500)) if i % 2 == 0 else int(r.integers(8, 300))
        kind = serving.KIND_LU if i % 3 else serving.KIND_CHOL
        a, b = system(r, kind, n, 1)
        specs.append((kind, a, b, 0.0005 * i))
    loop = serving.ServingLoop(lib, torch, dev, slots=4, mixed=mixed)
    mix = [request(torch, dev, i, k, a, b, t) for i, (k, a, b, t) in enumerate(specs)]
    torch.cuda.synchronize()
    t0 = __import__("time").perf_counter()
    loop.run(mix)
    t_mix = __import__("time").perf_counter() - t0
    solo = serving.ServingLoop(lib, torch, dev, slots=1, mixed=mixed)
    t_solo = 0.0
    for (kind, a, b, _), q in zip(specs, mix):
        assert q.status == 0 and q.eta <= 1e-13, (q.rid, q.n, q.path)
        alone = request(torch, dev, q.rid, kind, a, b)
        torch.cuda.synchronize()
        t0 = __import__("time").perf_counter()
        solo.run([alone])
        t_solo += __import__("time").perf_counter() - t0
        assert torch.equal(alone.x, q.x), (q.rid, q.n, q.path)
    print(f"40 requests: {t_mix * 1e3:.1f} ms in the mix, {t_solo * 1e3:.1f} ms one after another")
    assert t_mix < t_solo
    solo.close()
    loop.close()


def test_resident_factor_reuse_and_eviction(env):
    lib, torch, dev = env
    r = wl.rng(83)
    a, b1 = system(r, serving.KIND_LU, 300, 1)
    b2 = r.standard_normal((3, 300))
    loop = serving.ServingLoop(lib, torch, dev)
    first = request(torch, dev, 0, serving.KIND_LU, a, b1, keep_id=7)
    again = request(torch, dev, 1, serving.KIND_LU, a, b2, arrival=0.002, reuse_id=7)
    loop.run([first, again])
    assert first.status == 0 and again.status == 0 and again.path == "reuse"
    x2 = again.x.cpu().numpy().reshape(3, 300)
    for i in range(3):
        assert host_eta(a, x2[i], b2[i]) <= 1e-13
    loop.evict(7)
    late = request(torch, dev, 2, serving.KIND_LU, a, b1, reuse_id=7)
    loop.run([late])
    assert late.status == 1  # InvalidArg: the factor is no longer resident
    loop.close()


def test_admission_budget(env):
    """A request that can never fit the budget fails with RLG_E_OOM instead
    of blocking the queue; smaller requests behind it are still served."""
    lib, torch, dev = env
    r = wl.rng(87)
    big_a, big_b = system(r, serving.KIND_LU, 1100, 1)
    small_a, small_b = system(r, serving.KIND_LU, 64, 1)
    loop = serving.ServingLoop(lib, torch, dev, mem_budget_bytes=8 << 20)
    big = request(torch, dev, 0, serving.KIND_LU, big_a, big_b)
    small = request(torch, dev, 1, serving.KIND_LU, small_a, small_b, arrival=0.001)
    loop.run([big, small])
    assert big.status == serving.RLG_E_OOM
    assert small.status == 0 and small.eta <= 1e-13
    loop.close()


@pytest.mark.parametrize("small_mixed", [False, True], ids=["small_fp64", "small_mixed"])
def test_mixed_mode(env, small_mixed):
    """Mixed mode: the one-shot LU above order 512 takes the mixed solve; the
    small systems take the FP64 batched launch, or the mixed batched launch
    with `small_mixed`; every result meets the FP64 backward-error bound."""
    lib, torch, dev = env
    r = wl.rng(89)
    specs = [(serving.KIND_LU, wl.conditioned(600, 1e3, 5, 6), r.standard_normal((1, 600)))]
    for n in (40, 200, 512):
        a, b = system(r, serving.KIND_LU, n, 2)
        specs.append((serving.KIND_LU, a, b))
    a, b = system(r, serving.KIND_CHOL, 150, 1)
    specs.append((serving.KIND_CHOL, a, b))
    loop = serving.ServingLoop(lib, torch, dev, mixed=True, small_mixed=small_mixed)
    reqs = [request(torch, dev, i, k, a, b) for i, (k, a, b) in enumerate(specs)]
    loop.run(reqs)
    torch.cuda.synchronize()
    assert reqs[0].path == "single_mixed"
    small = "batched_mixed" if small_mixed else "batched"
    assert [q.path for q in reqs if q.n <= 512] == [small] * 4
    for (kind, a, b), q in zip(specs, reqs):
        assert q.status == 0 and q.eta <= 1e-13
        x = q.x.cpu().numpy().reshape(b.shape)
        for i in range(b.shape[0]):
            assert host_eta(a, x[i], b[i]) <= 1e-13
    loop.close()
