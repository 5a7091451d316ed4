"""Replica batches: independent P3BCA systems of uneven size in one sweep.

Each replica must behave exactly as its own single-system run (its own
bbox-rule periodic domain, its own counts):
  * exact path: every replica's state and per-step counts bit-identical to
    the same replica run alone, over 4 steps of tb3_batch_run;
  * FP64 / FP32-mixed: per-replica counts identical over 4 steps; after 2
    steps the state within 1e-12 of the array scale (FP64), or within 1e-4
    per particle with the RMS as floor (FP32-mixed, whose FP32 rounding turns
    last-bit differences into FP32-ulp jumps);
  * per-replica energies against the oracle, each replica on its own domain;
  * empty and tiny replicas, and a dense replica whose neighbourhoods
    overflow the cell kernel's staging (its owners go to the dense-owner
    kernel).
"""
import sys

import numpy as np
import pytest

import threebodymdalgorithms_amd as tb3
from conftest import uniform_particles

pytestmark = pytest.mark.gpu

G = np.array([0.0, -9.81, 0.0])


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def replicas(dense=False):
    """Uneven uniform replicas (density ~2, cutoff 2), plus tiny ones."""
    sizes = [700, 0, 1300, 2, 450, 1, 900]
    reps = []
    for i, n in enumerate(sizes):
        box = max((n / 2.0) ** (1 / 3), 1.0)
        reps.append(uniform_particles(n, box, 200 + i) if n else np.zeros((0, 3)))
    if dense:
        rng = np.random.default_rng(9)
        reps.append(rng.normal(0.0, 0.8, (600, 3)) + 5.0)
    return reps


def offsets(reps):
    return np.concatenate([[0], np.cumsum([len(r) for r in reps])]).astype(np.int64)


@pytest.fixture(scope="module")
def eng():
    e = tb3.Engine(0)
    yield e
    e.close()


def run_alone(eng, P, precision, steps):
    eng.set_particles(P, np.zeros_like(P), np.zeros_like(P), np.ones(len(P)))
    eng.configure("p3bca", cutoff=2.0, precision=precision)
    counts = [eng.step(0.001, G)[1] for _ in range(steps)]
    return eng.get_particles(), counts


def run_batch(eng, reps, precision, steps):
    P = np.concatenate(reps)
    eng.set_particles(P, np.zeros_like(P), np.zeros_like(P), np.ones(len(P)))
    eng.configure("p3bca", cutoff=2.0, precision=precision)
    eng.batch_set(offsets(reps))
    counts = eng.batch_run(steps, 0.001, G)
    return eng.get_particles(), counts


@pytest.mark.parametrize("precision", ["exact", "fp64", "fp32"])
@pytest.mark.parametrize("dense", [False, True])
def test_batch_equals_replicas_alone(eng, precision, dense):
    reps = replicas(dense)
    off = offsets(reps)
    exact = precision == "exact"
    steps = 4 if exact else 2
    say(f"batch {precision} dense={dense} sizes={[len(r) for r in reps]}")
    st, counts = run_batch(eng, reps, precision, steps)
    assert counts.shape == (steps, len(reps))
    if not exact:  # counts over 4 steps from a second batch run
        _, counts4 = run_batch(eng, reps, precision, 4)
    for r, P in enumerate(reps):
        if len(P) == 0:
            assert np.all(counts[:, r] == 0)
            continue
        alone, calone = run_alone(eng, P, precision, steps if exact else 4)
        sl = slice(off[r], off[r + 1])
        if exact:
            assert [int(c) for c in counts[:, r]] == calone, r
            for name in ("pos", "vel", "acc", "f"):
                assert np.array_equal(st[name][sl], alone[name]), (r, name)
        else:
            assert [int(c) for c in counts4[:, r]] == calone, r
            alone2, _ = run_alone(eng, P, precision, steps)
            for name in ("pos", "vel", "acc", "f"):
                if precision == "fp32":
                    # FP32 rounding is discontinuous in its FP64 inputs: the
                    # last-bit order difference of step 0 moves step 1's
                    # triplets by FP32 ulps, amplified where forces cancel,
                    # so the per-particle error is floored at the RMS value
                    num = np.linalg.norm(st[name][sl] - alone2[name], axis=1)
                    den = np.linalg.norm(alone2[name], axis=1)
                    assert (num / np.maximum(den, np.sqrt(np.mean(den ** 2)) + 1e-300)).max() <= 1e-4, (r, name)
                else:
                    scale = np.abs(alone2[name]).max() + 1e-300
                    assert np.abs(st[name][sl] - alone2[name]).max() <= 1e-12 * scale, (r, name)


def test_batch_energies(eng, oracle):
    reps = replicas()
    P = np.concatenate(reps)
    eng.set_particles(P)
    eng.configure("p3bca", nu=1.3, cutoff=2.0, precision="fp64")
    eng.batch_set(offsets(reps))
    say("batch energies")
    E, counts = eng.batch_energy()
    for r, Q in enumerate(reps):
        if len(Q) < 3:
            assert E[r] == 0.0 and counts[r] == 0
            continue
        L, _ = oracle.domain_size(Q)
        Eo, co = oracle.energy(Q, nu=1.3, cutoff=2.0, L=L)
        assert int(counts[r]) == co, r
        assert E[r] == pytest.approx(Eo, rel=1e-12), r


def test_batch_rejects_bad_offsets(eng):
    P = uniform_particles(100, 4.0, 3)
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=2.0, precision="fp64")
    with pytest.raises(tb3.TB3Error):
        eng.batch_set([0, 60, 50, 100])
    with pytest.raises(tb3.TB3Error):
        eng.batch_set([0, 50, 99])
    eng.configure("auta", precision="fp64")
    with pytest.raises(tb3.TB3Error):
        eng.batch_set([0, 50, 100])
