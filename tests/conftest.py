import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def rel_err(a, ref):
    """max |a - ref| / rms(ref): the tolerance metric of the float tests."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    return float(np.max(np.abs(a - ref)) / rms) if rms > 0 else float(np.max(np.abs(a - ref)))


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU step so a failing box still tells which test ran
    if request.node.get_closest_marker("gpu") is not None:
        print("\n[gpu-case] %s" % request.node.nodeid, flush=True)
    yield
