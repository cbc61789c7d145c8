import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def relative_gap(a, b):
    """tests/iterative.rs:21-32: largest |a - b| over the largest |b| (absolute if b is 0)."""
    import numpy as np
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    assert a.shape == b.shape
    scale = max(np.abs(b).max(initial=0.0), np.finfo(float).tiny)
    scale = 1.0 if scale < 1e-300 else scale
    return float(np.abs(a - b).max(initial=0.0) / scale)


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="session")
def gpu():
    import maille_amd
    if maille_amd.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return maille_amd
