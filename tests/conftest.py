import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "imod-rs_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtiltgpu's C ABI)")


@pytest.fixture(scope="session")
def gpu():
    import tiltgpu
    if tiltgpu.device_count() < 1:
        pytest.fail("no GPU visible (tests marked gpu need an MI355X)")
    return tiltgpu


def rel_err(a, b):
    """max |a-b| / RMS(b): the parity measure (bar: <= 1e-4)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    rms = np.sqrt(np.mean(b * b))
    return float(np.max(np.abs(a - b)) / (rms if rms > 0 else 1.0))


@pytest.fixture(scope="session")
def rel():
    return rel_err
