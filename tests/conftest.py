import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcubismnova_amd.so on the GPU)")


def _gpu_count():
    from cubismnova_amd import _abi
    import ctypes
    n = ctypes.c_int(0)
    rc = _abi.lib().cnv_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests fail loudly if the library or device is missing."""
    from cubismnova_amd import _abi
    _abi.lib()  # raises ImportError when the HIP library is not built
    n = _gpu_count()
    assert n >= 1, "no HIP device visible to libcubismnova_amd.so"
    return n
