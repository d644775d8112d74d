import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels; needs an MI355X (gfx950)")


@pytest.fixture(scope="session")
def mco():
    from oracle import mco as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def mcg():
    from minicore_amd import _lib
    _lib.load()
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libmcg")
    return _lib


@pytest.fixture(scope="session")
def ctx(mcg):
    """Shared context with the E-step tuning off, so the small parity cases run the
    BF16 MFMA kernels (tuning would pick the equivalent exact kernel for them)."""
    c = mcg.Context(0)
    c.tuning(False)
    yield c
    c.close()
