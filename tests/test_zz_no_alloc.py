"""GPU: rlg_lu_factor_f64_noalloc (the LU entry point with its scratch in a
static device array) against rlg_lu_decompose_f64_dev, the existing entry
point it shares its kernels with.  Bar: bit-identical packed factor and
pivots, info = n, on two calls in a row (the second one smaller, so it runs
on the scratch the first one left behind)."""
import ctypes as C

import numpy as np
import pytest

from rulinalg_amd import _lib

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    lib = _lib.load()
    assert lib.rlg_set_device(0) == 0, _lib.last_error()
    return lib


def reference_entry(lib, a0):
    """Packed factor and step pivots from rlg_lu_decompose_f64_dev."""
    n = a0.shape[0]
    a = a0.clone()
    torch.cuda.synchronize()
    h = C.c_void_p()
    assert lib.rlg_lu_decompose_f64_dev(n, a.data_ptr(), n, None, C.byref(h)) == 0, _lib.last_error()
    ipiv = np.empty(n, dtype=np.int64)
    assert lib.rlg_lu_ipiv(h, ipiv.ctypes.data) == 0
    lib.rlg_lu_free(h)
    assert lib.rlg_synchronize() == 0
    return a.cpu().numpy(), ipiv


@pytest.mark.parametrize("orders", [(300, 200), (1100, 520)])
def test_noalloc_matches_existing_entry(lib, orders):
    for call, n in enumerate(orders):
        print(f"noalloc call {call}: n={n}", flush=True)
        g = torch.Generator(device="cuda:0")
        g.manual_seed(700 + n)
        a0 = torch.randn(n, n, dtype=torch.float64, device="cuda:0", generator=g)
        a = a0.clone()
        ipiv = torch.full((n,), -7, dtype=torch.int64, device="cuda:0")
        info = torch.full((1,), -7, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        st = lib.rlg_lu_factor_f64_noalloc(n, a.data_ptr(), n, ipiv.data_ptr(), info.data_ptr(), None)
        assert st == 0, _lib.last_error()
        assert lib.rlg_synchronize() == 0
        assert int(info.item()) == n
        lu_ref, ipiv_ref = reference_entry(lib, a0)
        assert np.array_equal(ipiv.cpu().numpy(), ipiv_ref)
        assert np.array_equal(a.cpu().numpy(), lu_ref)


def test_noalloc_reports_singular_step(lib):
    n = 90
    a = torch.randn(n, n, dtype=torch.float64, device="cuda:0")
    a[:, 37] = 0.0
    ipiv = torch.empty(n, dtype=torch.int64, device="cuda:0")
    info = torch.empty(1, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    assert lib.rlg_lu_factor_f64_noalloc(n, a.data_ptr(), n, ipiv.data_ptr(), info.data_ptr(), None) == 0
    assert lib.rlg_synchronize() == 0
    assert int(info.item()) == 37


def test_noalloc_rejects_bad_operands(lib):
    info = torch.empty(1, dtype=torch.int64, device="cuda:0")
    assert lib.rlg_lu_factor_f64_noalloc(10, None, 10, None, info.data_ptr(), None) == 1
    assert lib.rlg_lu_factor_f64_noalloc(10, 1, 5, 1, info.data_ptr(), None) == 1
