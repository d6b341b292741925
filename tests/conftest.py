"""Shared test setup.

Markers: ``gpu`` tests need a HIP device and libmpsq.so; everything else
runs on the CPU (oracle pinning, host logic, ABI exports, error paths).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmpsq.so")
    config.addinivalue_line("markers", "slow: full-size property checks (seconds on the GPU)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    """Build the CPU oracle (gcc) if it is not there yet."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "oracle.c")
    if not os.path.exists(so) or os.path.getmtime(src) > os.path.getmtime(so):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "liboracle.so"])
    yield


@pytest.fixture(scope="session")
def gpu_lib():
    """libmpsq.so loaded; the GPU tests fail loudly if it is missing."""
    from mp_tools_amd import _lib
    return _lib.load()


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
