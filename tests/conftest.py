"""Shared pytest configuration: the gpu marker and helpers.

Tests marked gpu need a real device and the in-tree libself_amd.so; every
other test runs on the CPU (oracle checks, host set-up, C-ABI exports).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with device 0 selected; fails loudly when absent."""
    from self_amd import _lib as L
    n = __import__("ctypes").c_int()
    L.check(L.lib().self_device_count(__import__("ctypes").byref(n)))
    assert n.value >= 1, "no HIP device visible"
    L.check(L.lib().self_init(0))
    return L
