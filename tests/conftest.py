import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library with a visible GPU. GPU tests fail (not skip) without one:
    they only run where a GPU is expected."""
    from tiralib_amd import _lib
    L = _lib.load()
    n = _lib.device_count()
    assert n >= 1, "no GPU visible to libtira_hip.so"
    return _lib


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def rel_frob(a, b):
    """||a - b||_F / ||b||_F over the compared slots."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
