import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle32():
    from oracle.oracle import Oracle
    o = Oracle(32)
    o.set_threads(1)
    return o


@pytest.fixture(scope="session")
def oracle64():
    from oracle.oracle import Oracle
    o = Oracle(64)
    o.set_threads(1)
    return o


def _reference(bits):
    from oracle.oracle import Reference, reference_available, Oracle
    if not reference_available(bits):
        pytest.skip("reference C build (oracle/_ref) not present")
    Oracle(bits).set_threads(1)   # one OpenMP thread for bit-exact comparisons
    return Reference(bits)


@pytest.fixture(scope="session")
def reference32():
    return _reference(32)


@pytest.fixture(scope="session")
def reference64():
    return _reference(64)


def sphere(n, radius, inside=1.1, outside=1.0, centre=None):
    c = n // 2 if centre is None else centre
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    ball = r <= radius
    return np.where(ball, inside, outside).astype(np.float32), ball


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library through its C ABI; fails (not skips) if it is not built."""
    from disptools_amd import native
    if not native.set_device(0):
        pytest.fail("no HIP device visible")
    return native
