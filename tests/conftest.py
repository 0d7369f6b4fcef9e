import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)     # test helpers (fake_system)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture
def gue20():
    """20 x 20 GUE Hamiltonian (kwant.rmt.gaussian(20) recipe), fixed seed."""
    from oracle.brute import gaussian_unitary
    return gaussian_unitary(20, 1234)


def rel_err(x, y, scale):
    """max |x - y| / scale."""
    return float(np.max(np.abs(np.asarray(x) - np.asarray(y))) / scale)
