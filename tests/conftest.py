"""Shared test set-up: repo root on sys.path, the `gpu` marker, helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    if den == 0:
        return 0.0 if num == 0 else float("inf")
    return float(num / den)


@pytest.fixture(scope="session")
def gpu_available():
    import xfdtd_cuda_amd as xf
    if xf.device_count() < 1:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X")
    return True
