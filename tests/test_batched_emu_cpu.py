"""CPU: the batched solver's kernels, run from their own source
(rulinalg_amd/csrc/batched_device.hpp) under the emulation in tests/emu,
with AddressSanitizer and UndefinedBehaviorSanitizer.

Each of a workgroup's 256 lanes is a fiber.  Barriers and wave collectives
(lane reads, DPP moves, ballots, MFMA) are checked for divergence, and every
buffer the kernels touch is its own exact-size allocation, so any out-of-bounds
global access aborts the run.  LDS does not fault on the device (out-of-range
LDS accesses are dropped by the hardware), so a device memory fault in these
kernels would have to come from a global access.  The results are checked
against the oracle with the tolerances of the GPU batched tests.  Runs in
different lane orders must agree bit for bit, which rules out races between
a workgroup's threads on LDS or on the system's own memory.

Shapes: the autotuner's (LU of orders 24, 56, 120, 250, 500 with 4 right-hand
sides, at each panel width it may pick: 8, 16, and 32 where
(n + 16) * 33 doubles fit the LDS panel, batched.hip `nb_for`), Cholesky at
the same shapes, edge orders 1..512 with 1 and 16 right-hand sides in one
persistent launch, the mixed form in the last workspace slot of a 256-slot
grid, its FP64 fallback, and the singular / indefinite failure paths.
"""
import os
import struct
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as orc
from rulinalg_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
EXE = os.path.join(EMU, "_build", "batched_emu")
SELFTEST = os.path.join(EMU, "_build", "emu_selftest")
CXX = "/opt/rocm/llvm/bin/clang++"
LU, CHOL = 0, 1
OK, DECOMP_FAILURE, DIV_BY_ZERO = 0, 2, 4  # rlg_status in include/rlgpu.h
PANEL = 8704  # doubles in the kernels' LDS panel
U = np.finfo(np.float64).eps
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")

pytestmark = pytest.mark.skipif(not os.path.exists(CXX), reason="no clang++ for the emulation build")


@pytest.fixture(scope="module", autouse=True)
def build():
    p = subprocess.run(["make", "-s", "-j4", "-C", EMU], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]


def run(systems, mode=0, block=0, grid=1, order=0, seed=1, max_iter=30, tol=1e-13):
    """Runs (kind, a, b, nb) systems in one emulated launch; returns
    (statuses, fallbacks, [(factor, solution, ipiv)])."""
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<7i", mode, len(systems), block, grid, order, seed, max_iter))
            f.write(struct.pack("<d", tol))
            for kind, a, b, nb in systems:
                f.write(struct.pack("<4i", kind, a.shape[0], b.shape[0], nb))
            for kind, a, b, nb in systems:
                f.write(np.ascontiguousarray(a, "<f8").tobytes())
                f.write(np.ascontiguousarray(b, "<f8").tobytes())
        p = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=600, env=ENV)
        assert p.returncode == 0, f"emulated launch failed ({p.returncode}): {p.stderr[-3000:]}"
        raw = open(fout, "rb").read()
    c = len(systems)
    st = list(struct.unpack_from(f"<{c}i", raw, 0))
    off = 4 * c
    fb, = struct.unpack_from("<i", raw, off)
    off += 4 + 16
    out = []
    for kind, a, b, nb in systems:
        n, r = a.shape[0], b.shape[0]
        fa = np.frombuffer(raw, "<f8", n * n, off).reshape(n, n)
        off += 8 * n * n
        x = np.frombuffer(raw, "<f8", r * n, off).reshape(r, n)
        off += 8 * r * n
        ip = np.frombuffer(raw, "<i8", n, off)
        off += 8 * n
        out.append((fa, x, ip))
    return st, fb, out


def eta(a, x, b):
    return np.abs(b - a @ x).max() / (np.abs(a).sum(1).max() * np.abs(x).max() + np.abs(b).max())


def check(kind, a, b, fa, x, ip, factor_tol=1e-10):
    """The GPU batched tests' rules: pivots equal to the oracle's where its
    margins are clear, backward error <= 1e-13, forward difference <=
    max(1e-13, 16 n kappa u), Cholesky factor within `factor_tol`."""
    n = a.shape[0]
    k = np.linalg.cond(a, np.inf)
    if kind == LU:
        lu_or, perm_or, ipiv_or, margin = orc.lu_decompose(a, with_margin=True)
        clear = margin > 1e-9 * np.abs(a).max()
        fu = int(np.argmin(clear)) if not np.all(clear) else n
        assert np.array_equal(ip[:fu], ipiv_or[:fu])
        ref = [orc.lu_solve(lu_or, perm_or, b[i]) for i in range(b.shape[0])]
    else:
        w = orc.chol_decompose(a)
        l_or = orc.chol_unpack(w)
        assert np.abs(np.tril(fa) - l_or).max() / np.abs(l_or).max() <= factor_tol
        ref = [orc.chol_solve(w, b[i]) for i in range(b.shape[0])]
    for i in range(b.shape[0]):
        assert eta(a, x[i], b[i]) <= 1e-13
        assert np.abs(x[i] - ref[i]).max() / np.abs(ref[i]).max() <= max(1e-13, 16 * n * k * U)


def system(r, kind, n, nrhs):
    if kind == LU:
        a = r.standard_normal((n, n))
    else:
        g = r.standard_normal((n, n))
        a = g @ g.T / n + np.eye(n)
    return a, r.standard_normal((nrhs, n))


def widths(n):
    return [8, 16] + ([32] if (n + 16) * 33 <= PANEL else [])


def parallel(fn, cases):
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        return list(ex.map(fn, cases))


AUTOTUNE_N = [24, 56, 120, 250, 500]


@pytest.mark.parametrize("kind", [LU, CHOL], ids=["lu", "chol"])
def test_autotuner_shapes_plain(kind):
    cases = [(n, nb) for n in AUTOTUNE_N for nb in widths(n)]

    def one(case):
        n, nb = case
        r = wl.rng(1000 + 7 * n + nb)
        a, b = system(r, kind, n, 4)
        st, fb, [(fa, x, ip)] = run([(kind, a, b, nb)])
        return case, a, b, st, fa, x, ip

    for (n, nb), a, b, st, fa, x, ip in parallel(one, cases):
        assert st == [OK], (n, nb, st)
        check(kind, a, b, fa, x, ip)


def test_edge_orders_in_one_launch():
    """Orders at and around every panel and tile boundary, both kinds, 1 and
    16 right-hand sides, in one persistent launch (the workgroup takes item
    after item from the counter)."""
    r = wl.rng(1100)
    orders = [1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 65, 127, 129, 255, 257, 511, 512]
    systems = []
    for i, n in enumerate(orders):
        kind = LU if i % 2 == 0 else CHOL
        nb = [8, 16, 32][i % 3]
        if nb == 32 and (n + 16) * 33 > PANEL:
            nb = 16
        a, b = system(r, kind, n, 16 if i % 4 < 2 else 1)
        systems.append((kind, a, b, nb))
    st, fb, out = run(systems)
    assert st == [OK] * len(systems)
    for (kind, a, b, nb), (fa, x, ip) in zip(systems, out):
        check(kind, a, b, fa, x, ip)


def test_results_independent_of_lane_order():
    """The same launch with the lanes run ascending, descending and shuffled
    (two seeds) between synchronization points: bit-identical factors,
    solutions and pivots, plain and mixed."""
    r = wl.rng(1200)
    systems = []
    for kind in (LU, CHOL):
        for n, nb in ((40, 8), (120, 16), (100, 32), (250, 16)):
            a, b = system(r, kind, n, 3)
            systems.append((kind, a, b, nb))
    runs = [(mode, order, seed) for mode in (0, 1) for order, seed in ((0, 1), (1, 1), (2, 1), (2, 2))]

    def one(cfg):
        mode, order, seed = cfg
        return cfg, run(systems, mode=mode, order=order, seed=seed, grid=2, block=1)

    res = dict(parallel(one, runs))
    for mode in (0, 1):
        base = res[(mode, 0, 1)]
        assert base[0] == [OK] * len(systems)
        for cfg in runs:
            if cfg[0] != mode or cfg == (mode, 0, 1):
                continue
            other = res[cfg]
            assert other[0] == base[0] and other[1] == base[1]
            for (f0, x0, p0), (f1, x1, p1) in zip(base[2], other[2]):
                assert np.array_equal(f0, f1) and np.array_equal(x0, x1) and np.array_equal(p0, p1), cfg


def test_mixed_autotuner_shapes_last_slot():
    """The mixed kernel at the autotuner's shapes, run as the last workgroup of
    a 256-workgroup grid (the last workspace slot): no fallback, backward
    error <= 1e-13, A left unmodified."""
    cases = [(kind, n, nb) for kind in (LU, CHOL) for n in AUTOTUNE_N for nb in widths(n)[:2]]

    def one(case):
        kind, n, nb = case
        r = wl.rng(1300 + 7 * n + nb + kind)
        a, b = system(r, kind, n, 4)
        st, fb, [(fa, x, ip)] = run([(kind, a, b, nb)], mode=1, grid=256, block=255)
        return case, a, b, st, fb, fa, x

    for case, a, b, st, fb, fa, x in parallel(one, cases):
        assert st == [OK] and fb == 0, (case, st, fb)
        assert np.array_equal(fa, a)
        for i in range(b.shape[0]):
            assert eta(a, x[i], b[i]) <= 1e-13, case


def test_mixed_fallback_and_failure_paths():
    """A system too ill-conditioned for FP32 refinement takes the FP64
    fallback inside the kernel and still solves to 1e-13; a singular LU
    reports RLG_E_DIV_BY_ZERO and an indefinite Cholesky
    RLG_E_DECOMP_FAILURE, in both forms, without an emulation error."""
    r = wl.rng(1400)
    n = 96
    u, _ = np.linalg.qr(r.standard_normal((n, n)))
    v, _ = np.linalg.qr(r.standard_normal((n, n)))
    hard = (u * np.logspace(0, -9, n)) @ v.T  # kappa 1e9: FP32 refinement stalls
    b_hard = r.standard_normal((2, n))
    sing = r.standard_normal((64, 64))
    sing[:, 10] = 0.0
    indef = np.eye(48)
    indef[20, 20] = -1.0
    systems = [(LU, hard, b_hard, 16), (LU, sing, r.standard_normal((1, 64)), 8),
               (CHOL, indef, r.standard_normal((1, 48)), 16)]
    st, fb, out = run(systems, mode=1, grid=4, block=3)
    assert st == [OK, DIV_BY_ZERO, DECOMP_FAILURE]
    assert fb == 3  # every system that does not converge in FP32 is refactored in FP64
    for i in range(2):
        assert eta(hard, out[0][1][i], b_hard[i]) <= 1e-13
    st, fb, out = run(systems[1:])
    assert st == [DIV_BY_ZERO, DECOMP_FAILURE]


def test_random_mix_in_one_launch():
    """A seeded random mix in one launch of each form: orders 1-512, 1-16
    right-hand sides, both kinds, every panel width the host allows for the
    order; each item against the oracle (plain) or to backward error 1e-13
    with A left unmodified (mixed)."""
    r = wl.rng(1500)
    systems = []
    for i in range(24):
        n = int(r.integers(1, 513))
        kind = LU if r.random() < 0.5 else CHOL
        nb = int(r.choice(widths(n)))
        a, b = system(r, kind, n, int(r.integers(1, 17)))
        systems.append((kind, a, b, nb))
    st, fb, out = run(systems, order=2, seed=5)
    assert st == [OK] * len(systems)
    for (kind, a, b, nb), (fa, x, ip) in zip(systems, out):
        check(kind, a, b, fa, x, ip)
    st, fb, out = run(systems, mode=1, grid=8, block=7)
    assert st == [OK] * len(systems) and fb == 0
    for (kind, a, b, nb), (fa, x, ip) in zip(systems, out):
        assert np.array_equal(fa, a)
        for i in range(b.shape[0]):
            assert eta(a, x[i], b[i]) <= 1e-13


@pytest.mark.parametrize("case, expect", [("clean", 0), ("barrier", 2), ("halfwave", 2), ("oob", 1)])
def test_emulation_reports_planted_faults(case, expect):
    """Planted in small kernels of tests/emu/emu_selftest.cpp: a barrier that
    only one wave reaches, a lane read by half a wave (emulation errors), a
    write past the end of a buffer (AddressSanitizer)."""
    p = subprocess.run([SELFTEST, case, "0"], capture_output=True, text=True, timeout=60, env=ENV)
    assert p.returncode == expect, p.stderr[-2000:]
    if case == "oob":
        assert "heap-buffer-overflow" in p.stderr


def test_emulation_exposes_a_planted_race():
    """A shared value read without the barrier after thread 0 writes it gives
    different results in ascending and descending lane order."""
    outs = [subprocess.run([SELFTEST, "race", o], capture_output=True, text=True, timeout=60, env=ENV)
            for o in ("0", "1")]
    assert all(p.returncode == 0 for p in outs)
    assert outs[0].stdout != outs[1].stdout
    clean = [subprocess.run([SELFTEST, "clean", o], capture_output=True, text=True, timeout=60, env=ENV).stdout
             for o in ("0", "1", "2")]
    assert clean[0] == clean[1] == clean[2]

