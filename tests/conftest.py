import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X")


@pytest.fixture(scope="session")
def golden_1000():
    return np.load(os.path.join(GOLDEN, "2X_1D_1000_4.npy")).astype(np.float32)


@pytest.fixture(scope="session")
def golden_2000():
    return np.load(os.path.join(GOLDEN, "2X_1D_2000_4.npy")).astype(np.float32)


def curve_inputs(n, d, variant="mi"):
    """The deterministic curves of the reference's netcalc tests
    (tests/netcalc/python/test_MutualInformation.py:67-154,
    test_GeneralizedCorrelation.py:12-226): x starts at 0.001 and advances
    by pi/n in double, values stored as float32. Returns flat Xa, Xb of d*n."""
    xa = np.zeros((d, n), np.float32)
    xb = np.zeros((d, n), np.float32)
    x = 0.001
    for i in range(n):
        if variant == "mi":  # MI tests: Xa=(sin, cos), Xb=(x, 2x)
            rows_a = [np.sin(x), np.cos(x), np.log(x)]
            rows_b = [x, 2 * x, np.log(2 * x)]
        else:  # GC tests: Xa=(sin, x, log x), Xb=(cos, 2x, log 2x)
            rows_a = [np.sin(x), x, np.log(x)]
            rows_b = [np.cos(x), 2 * x, np.log(2 * x)]
        for di in range(d):
            xa[di, i] = float(rows_a[di])
            xb[di, i] = float(rows_b[di])
        x += float(np.pi / n)
    return xa.ravel(), xb.ravel()
