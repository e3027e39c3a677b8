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


def selection_weights(seed: int) -> np.ndarray:
    """fp16 weights of a learned layer whose matrix products are exact: W1
    rows and W2 rows each hold one non-zero entry (+-0.5, 1 or 2 on a random
    input / a permutation of the hidden units), biases b1 = b2 = 0, w3 and b3
    random.  Every 8-term group the matrix cores add then holds at most one
    non-zero term, so their sums are exact whatever order the hardware adds
    in, and the engine must match the oracle bit for bit; a wrong lane
    layout or hidden-unit order would not."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((32, 12))
    for o in range(32):
        W1[o, rng.integers(12)] = rng.choice([-2.0, -1.0, -0.5, 0.5, 1.0, 2.0])
    W2 = np.zeros((32, 32))
    for o, u in enumerate(rng.permutation(32)):
        W2[o, u] = rng.choice([-2.0, -1.0, -0.5, 0.5, 1.0, 2.0])
    w3 = rng.standard_normal(32) * 0.05
    w = np.concatenate([W1.ravel(), np.zeros(32), W2.ravel(), np.zeros(32), w3, [0.003]])
    return w.astype(np.float16)


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
