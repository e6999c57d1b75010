import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_golden(name):
    return np.loadtxt(os.path.join(GOLDEN, name))


@pytest.fixture(scope="session")
def lj38():
    return load_golden("LJ38_coords.txt"), load_golden("LJ38_finish.txt")


@pytest.fixture(scope="session")
def blj256():
    return load_golden("BLJ256_coords.txt"), load_golden("BLJ256_finish.txt")


def gpu_available():
    try:
        import ctypes
        from fastoverlap_amd import _lib
        n = ctypes.c_int(0)
        return _lib.load().fo_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails (not skips) when the HIP library or the device is missing: a GPU
    test must never pass on a fallback path."""
    from fastoverlap_amd import _lib
    _lib.load()
    assert gpu_available(), "no GPU visible to libfastoverlap_hip"
    return _lib.default_context()
