import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "ref")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~10 s")


def read_dat(name, dtype=np.float64, shape=None):
    a = np.fromfile(os.path.join(GOLDEN, name), dtype=dtype)
    return a.reshape(shape) if shape is not None else a


def l1rel(a, b):
    """L1-relative error sum|a-b| / sum|b| over complex entries (test/Cc_test.c:177-189)."""
    return float(np.abs(np.asarray(a) - np.asarray(b)).sum() / np.abs(np.asarray(b)).sum())


def rand_state(rng, ncell, N, L=6.0):
    """Seeded physical Wigner states: w0 a positive Gaussian bump, (w1,w2,w3) = 0.8 w0 n, |n| = 1."""
    v = np.array([k if k < N // 2 else k - N for k in range(N)]) * 2 * L / N
    vx, vy = np.meshgrid(v, v, indexing="ij")
    out = np.empty((ncell, 4, N, N))
    for c in range(ncell):
        cx_, cy_ = rng.uniform(-1.0, 1.0, 2)
        s = rng.uniform(0.8, 1.2)
        w0 = rng.uniform(0.2, 0.5) * np.exp(-((vx - cx_) ** 2 + (vy - cy_) ** 2) / (2 * s * s))
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        out[c, 0] = w0
        for i in range(3):
            out[c, i + 1] = 0.8 * w0 * n[i]
    return out
