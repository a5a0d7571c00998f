import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SDESIM_TEST_PACKAGE_ROOT: the directory holding the sde_sim_rs package under test (the sanitized
# build's copy, tests/test_sanitized_host.py); the in-tree package otherwise.
_PKG_ROOT = os.environ.get("SDESIM_TEST_PACKAGE_ROOT") or os.path.join(ROOT, "sde-sim-rs_amd")
for p in (ROOT, _PKG_ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


GBM = ["dX = (0.05 * X) * dt + (0.2 * X) * dW1"]
HESTON = [
    "dS = (0.02 * S) * dt + (max(v, 0)^0.5 * S) * dW1",
    "dv = (2.0 * (0.04 - max(v, 0))) * dt + (-0.7 * 0.5 * max(v, 0)^0.5) * dW1"
    " + (0.7141428428542850 * 0.5 * max(v, 0)^0.5) * dW2",
]
JUMP = [
    "dX0 = ( 2.0 * (0.5 - X0) ) * dt + ( 0.1 ) * dN1(X0)",
    "dX1 = ( 0.05 * X1 ) * dt + ( 0.2 * X1 ) * dW1 + ( 0.5 ) * dN1(X0)",
    "X2 = max(X1 - 100.0, 0.0)",
]
JUMP_X0 = {"X0": 0.5, "X1": 100.0, "X2": 0.0}


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0
