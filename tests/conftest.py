import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # The serving tests hold their device buffers in PyTorch tensors.  In a
    # process that uses both, PyTorch has to be loaded before librlgpu.so:
    # loaded after it, PyTorch finds no HIP device.  Runs that may select the
    # GPU tests import it here, before any test module loads the library.
    markexpr = config.getoption("markexpr", "") or ""
    if "not gpu" not in markexpr:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
