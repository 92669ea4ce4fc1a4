import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

INSTR = os.path.join(ROOT, "ramp_amd", "instruments")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libramp.so on the GPU)")


@pytest.fixture(scope="session")
def lib():
    from ramp_amd import _capi
    return _capi.load_library()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def gctx(lib):
    import ctypes as C
    from ramp_amd import _capi
    h = C.c_void_p()
    _capi.check(lib.ramp_ctx_create(0, C.byref(h)))
    yield h
    lib.ramp_ctx_destroy(h)
