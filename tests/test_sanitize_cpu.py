"""CPU: AddressSanitizer + UndefinedBehaviorSanitizer runs of the host code.

* the oracle (oracle/rlg_oracle.c) with a driver that runs every routine at
  the edge orders 0, 1, 2, 7, 8, 9, 16, 65 (`make -C oracle sanitize`);
* librlgpu's host half (every HIP source compiled with the sanitizers on its
  host code only) with a driver that calls every entry point of
  include/rlgpu.h through its argument checks and host-only paths, without a
  GPU (`make -C tests/sanitize`; about half a minute of compiling).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, timeout):
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, (p.stdout[-3000:] + p.stderr[-3000:])
    return p.stdout


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None, reason="no C compiler")
def test_oracle_under_asan_ubsan():
    out = _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], 300)
    assert "clean" in out


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="no hipcc")
def test_library_host_code_under_asan_ubsan():
    jobs = str(min(8, os.cpu_count() or 2))
    out = _run(["make", "-s", "-j" + jobs, "-C", os.path.join(ROOT, "tests", "sanitize")], 900)
    assert "clean" in out
