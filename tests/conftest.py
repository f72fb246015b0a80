import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libpawan_gpu.so")


def normwise(a, b):
    """max_i |a_i - b_i| / max_i |b_i| over particles (rows)."""
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    den = np.max(np.linalg.norm(b, axis=1))
    return float(np.max(np.linalg.norm(a - b, axis=1)) / den) if den > 0 else float(np.max(np.abs(a)))


def rmswise(a, b):
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O
