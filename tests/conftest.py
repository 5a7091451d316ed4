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


def force_longdouble(P, p, cutoff, L):
    """+grad E at particle p, summed over p's triplets, in 80-bit extended
    precision (numpy longdouble on x86-64) with the squared-distance form
    E = nu (1 + 3 cos cos cos) / (r r r)^3, nu = 1.  The triplet set is the
    reference's: every pair within the cutoff by the FP64 periodic test of
    structs.hpp:106-115 (the same operations in float64 here), or every pair
    when cutoff <= 0 (L is then unused).  A reference
    for the true force against which both FP64 formulations are measured."""
    P = np.asarray(P, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    c2 = cutoff * cutoff

    def d2(a, b):
        t = np.abs(b - a)
        e = np.minimum(t, L - t)
        return e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1] + e[..., 2] * e[..., 2]

    near = np.flatnonzero(~(d2(P[p], P) > c2)) if cutoff > 0 else np.arange(len(P))
    near = near[near != p]
    if len(near) < 2:
        return np.zeros(3)
    a, b = np.triu_indices(len(near), 1)
    q, r = near[a], near[b]
    if cutoff > 0:
        ok = ~(d2(P[q], P[r]) > c2)
        q, r = q[ok], r[ok]
    ld = np.longdouble
    xi = P[p].astype(ld)
    dj = P[q].astype(ld) - xi
    dk = P[r].astype(ld) - xi
    djk = P[r].astype(ld) - P[q].astype(ld)
    u, v, w = (dj * dj).sum(1), (dk * dk).sum(1), (djk * djk).sum(1)
    K = ld(0.375)
    s = 1 / np.sqrt(u * v * w)
    s2 = s * s
    A, B, C = u + v - w, u - v + w, -u + v + w
    X, Y, Z = B * C, A * C, A * B
    Q = A * X * s2
    R = -ld(1.5) - ld(2.5) * K * Q
    S = X + Y + Z
    s5 = s2 * s2 * s
    Eu = s5 * (R * v * w + K * (S - 2 * Z))
    Ev = s5 * (R * u * w + K * (S - 2 * Y))
    g = -2 * Eu[:, None] * dj - 2 * Ev[:, None] * dk
    return g.sum(0).astype(np.float64)
