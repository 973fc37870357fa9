import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

# torch ships its own HIP runtime under the same soname as the system one; it
# must be loaded before libhs_amd.so pulls in the system copy, or torch's
# device initialisation fails in a process that also uses the library.
try:
    import torch  # noqa: F401
except ImportError:
    pass

from hs_testlib import Oracle, load_hs  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and runs through libhs_amd.so")


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def hs():
    return load_hs()


@pytest.fixture(scope="session")
def dev(hs):
    import ctypes as C

    lib = hs.load()
    cnt = C.c_int(0)
    rc = lib.hs_device_count(C.byref(cnt))
    if rc != 0 or cnt.value == 0:
        pytest.skip("no GPU visible to the HIP runtime")
    # the largest systems the GPU tests solve on the shared handle: the
    # workspace and event pool are sized once here, as a caller would at
    # set-up (no solver entry allocates)
    d = hs.Device(0, reserve=[(65536, 64), (65536, 128), (65536, 256, 16384), (4096, 1)])
    yield d
    d.close()
