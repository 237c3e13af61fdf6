import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

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
