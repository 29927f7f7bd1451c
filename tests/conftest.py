import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


def rel_linf(a, b, nvars):
    """Per-variable relative L-infinity: max|a-b| / max|b| (absolute where max|b| == 0)."""
    a = np.asarray(a).reshape(nvars, -1)
    b = np.asarray(b).reshape(nvars, -1)
    out = []
    for k in range(nvars):
        den = np.max(np.abs(b[k]))
        num = np.max(np.abs(a[k] - b[k]))
        out.append(num / den if den > 0 else num)
    return max(out)
