import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
    config.addinivalue_line("markers", "slow: full-size configuration, minutes of CPU oracle time")


def csv_round(a):
    """Round values through the reference CSV text format (%g, 6 significant digits)."""
    flat = np.asarray(a, dtype=np.float64).ravel()
    return np.array([float("%g" % v) for v in flat]).reshape(np.shape(a))


def uniform_particles(n, box, seed):
    rng = np.random.default_rng(seed)
    return csv_round(rng.uniform(0.0, box, size=(n, 3)))


def clustered_particles(n, nclusters, box, sigma, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, box, size=(nclusters, 3))
    which = np.arange(n) % nclusters
    return csv_round(centres[which] + rng.normal(0.0, sigma, size=(n, 3)))


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib
