import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X) and libgravhip.so')
    config.addinivalue_line('markers', 'slow: CPU test that takes more than a few seconds')


@pytest.fixture(scope='session')
def cuda():
    """The HIP device; a GPU test fails (it does not skip) when none is visible."""
    import torch
    from gravitation_amd import _native
    _native.load()
    assert torch.cuda.is_available(), 'GPU test run without a visible HIP device'
    return torch.device('cuda:0')


def oracle_rows(pos, m, G, eps=0.0, rows=None):
    """FP64 oracle accelerations for the given rows: the C oracle when it is built, else the
    NumPy oracle."""
    from oracle import oracle as O, ref as R
    if R.available('oracle'):
        return R.oracle_accel_c(pos, m, G, eps, rows)
    return O.accel_f64(pos, m, G, eps, rows)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
