"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and run on an MI355X;
everything else runs on the CPU (oracle, host code, library symbol checks)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def rel_err(a, b):
    """(max, mean) relative error of a against b over vertices with finite b > 0;
    both must agree on which vertices are infinite."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert np.array_equal(np.isinf(a), np.isinf(b)), "reached sets differ"
    m = np.isfinite(b) & (b > 0)
    if not m.any():
        return 0.0, 0.0
    r = np.abs(a[m] - b[m]) / b[m]
    return float(r.max()), float(r.mean())


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
