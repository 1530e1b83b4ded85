import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~10 s")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when the HIP library or the device is missing (no CPU fallback)."""
    import ctypes as C

    from mycel_amd import _lib as L

    lib = L.load()
    n = C.c_int32()
    L.check(lib.mycel_device_count(C.byref(n)))
    assert n.value >= 1, "no HIP device visible"
    return lib
