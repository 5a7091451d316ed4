"""Multi-step runs (tb3_run) replay one captured step graph.  Each run is
checked against the same steps taken one tb3_step at a time, with eager work
and host synchronisation between the replays:
  * exact path: every state array and every count bit-identical;
  * FP64 path: counts identical, state within 1e-12 (its force sums are
    atomic, so their order is free);
  * runs in which a sweep overflows part way (two clusters that merge):
    the halted replays leave the state alone and the host finishes the step,
    so the result still equals the one-step-at-a-time run (bit for bit on the
    exact path), also where the dense-owner kernel is the one that overflows,
    and where a cell outgrows the slots of the slotted cell layout.
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


@pytest.fixture(scope="module")
def eng():
    e = tb3.Engine(0)
    yield e
    e.close()


def state_after(eng, P, V, algo, cutoff, precision, schedule, L=None, nu=1.0):
    """Runs the schedule (list of ("step", 1) / ("run", k)) from (P, V)."""
    eng.set_particles(P, V, np.zeros_like(P), np.ones(len(P)))
    eng.configure(algo, nu=nu, cutoff=cutoff, precision=precision, domain_L=L)
    counts = []
    for kind, k in schedule:
        if kind == "step":
            counts.append(eng.step(0.001, G)[1])
        else:
            counts.extend(int(c) for c in eng.run(k, 0.001, G))
    return eng.get_particles(), counts


def assert_same(a, ca, b, cb, exact, tol=1e-12, floored=False):
    assert ca == cb
    for name in ("pos", "vel", "acc", "f"):
        if exact:
            assert np.array_equal(a[name], b[name]), name
        elif floored:
            num = np.linalg.norm(a[name] - b[name], axis=1)
            den = np.linalg.norm(b[name], axis=1)
            assert (num / np.maximum(den, np.sqrt(np.mean(den ** 2)) + 1e-300)).max() <= tol, name
        else:
            scale = np.abs(b[name]).max() + 1e-300
            assert np.abs(a[name] - b[name]).max() <= tol * scale, name


@pytest.mark.parametrize("precision", ["exact", "fp64", "fp32"])
def test_run_equals_steps_p3bca(eng, precision):
    """Exact: 12 steps bit for bit.  FP64 / FP32: the atomic force sums make
    the trajectory order-dependent at the last bit and the dense random start
    amplifies that quickly, so the state is compared after 2 steps and the
    counts over 12 (FP64) or 4 (FP32, whose per-triplet rounding is larger:
    by step 7 a triplet or two cross the cutoff differently)."""
    n = 3000
    P = uniform_particles(n, (n / 2.0) ** (1 / 3), 41)
    V = np.zeros_like(P)
    exact = precision == "exact"
    say(f"run vs steps p3bca {precision}")
    ncmp = 4 if precision == "fp32" else 12
    ref, cref = state_after(eng, P, V, "p3bca", 2.0, precision, [("step", 1)] * 12)
    got, cgot = state_after(eng, P, V, "p3bca", 2.0, precision, [("run", 5), ("step", 1), ("run", 6)])
    assert cgot[:ncmp] == cref[:ncmp]
    if exact:
        assert_same(got, cgot, ref, cref, True)
    # the captured graph is reused by a second run on a new state
    got2, cgot2 = state_after(eng, P, V, "p3bca", 2.0, precision, [("run", 12)])
    assert cgot2[:ncmp] == cref[:ncmp]
    # the fast paths' replays bin into the slotted cell layout, the exact path's do not
    assert (eng.get_cell_slots()[1] > 0) == (not exact)
    if exact:
        assert_same(got2, cgot2, ref, cref, True)
    else:
        ref2, cref2 = state_after(eng, P, V, "p3bca", 2.0, precision, [("step", 1)] * 2)
        got3, cgot3 = state_after(eng, P, V, "p3bca", 2.0, precision, [("run", 2)])
        # FP32 rounding of u, v, w is discontinuous in the FP64 inputs: an
        # order-dependent last bit in step 0's forces can move a triplet's
        # FP32 result by an FP32 ulp at step 1 (amplified where forces
        # cancel), so FP32 runs are compared per particle with the RMS as
        # floor at 1e-4, FP64 runs at 1e-12 of the array scale
        if precision == "fp32":
            assert_same(got3, cgot3, ref2, cref2, False, tol=1e-4, floored=True)
        else:
            assert_same(got3, cgot3, ref2, cref2, False)


def test_run_uses_every_batched_graph(eng):
    """A run of 32 steps replays its 30 middle steps as batched graphs of 16,
    8, 4 and 2 steps: on the exact path the run equals 32 single steps bit
    for bit, so every batched graph replays the same step."""
    P = uniform_particles(1500, (1500 / 2.0) ** (1 / 3), 46)
    V = np.zeros_like(P)
    say("run of 32 steps vs 32 steps, exact")
    ref, cref = state_after(eng, P, V, "p3bca", 2.0, "exact", [("step", 1)] * 32)
    got, cgot = state_after(eng, P, V, "p3bca", 2.0, "exact", [("run", 32)])
    assert_same(got, cgot, ref, cref, True)


def test_run_equals_steps_auta(eng):
    P = uniform_particles(300, 5.0, 43)
    V = np.zeros_like(P)
    say("run vs steps auta exact")
    ref, cref = state_after(eng, P, V, "auta", 0.0, "exact", [("step", 1)] * 4)
    got, cgot = state_after(eng, P, V, "auta", 0.0, "exact", [("run", 4)])
    assert_same(got, cgot, ref, cref, True)


def merging_clusters():
    rng = np.random.default_rng(8)
    a = rng.normal(0.0, 0.25, (110, 3)) + np.array([6.0, 10.0, 10.0])
    b = rng.normal(0.0, 0.25, (110, 3)) + np.array([14.0, 10.0, 10.0])
    P = np.concatenate([a, b])
    V = np.zeros_like(P)
    V[:110, 0] = 600.0   # 0.6 per step at dt = 0.001: the clusters overlap at step 5
    V[110:, 0] = -600.0
    return P, V


@pytest.mark.parametrize("precision", ["exact", "fp64"])
def test_run_with_overflow_midway(precision):
    """nu = 1e-9 keeps the clusters intact while they drift into each other;
    each schedule starts on a new engine so both meet the overflow."""
    P, V = merging_clusters()
    L = np.array([20.0, 20.0, 20.0])
    say(f"run with a mid-run overflow {precision}")
    with tb3.Engine(0) as e1:
        ref, cref = state_after(e1, P, V, "p3bca", 2.0, precision, [("step", 1)] * 9, L=L, nu=1e-9)
    with tb3.Engine(0) as e2:
        got, cgot = state_after(e2, P, V, "p3bca", 2.0, precision, [("run", 9)], L=L, nu=1e-9)
    assert max(cref) > 3 * cref[0]  # the neighbourhoods did grow
    assert_same(got, cgot, ref, cref, precision == "exact")


def slot_lattice():
    """8 particles in each of the 6^3 cells (width 2.1: a 12.6^3 box, cutoff
    2) : the particles at x offset 1.55 of the cells with even x index drift
    at 100 / s, so that from step 5 on those cells' right neighbours hold 12."""
    rng = np.random.default_rng(23)
    off = np.array([[a, b, c] for a in (0.5, 1.55) for b in (0.5, 1.5) for c in (0.5, 1.5)])
    cells = np.array([[i, j, k] for i in range(6) for j in range(6) for k in range(6)], dtype=float)
    P = (2.1 * cells[:, None, :] + off[None, :, :]).reshape(-1, 3)
    P += rng.uniform(-0.02, 0.02, P.shape)
    V = np.zeros_like(P)
    drift = (np.repeat(cells[:, 0], 8) % 2 == 0) & (np.tile(off[:, 0], len(cells)) == 1.55)
    V[drift, 0] = 100.0
    return P, V


def test_run_slots_overflow_midway(eng):
    """With 8 slots per cell the steps replay on the slotted layout until a
    cell receives more particles than slots (step 5); that step and the rest
    of the run then take the contiguous layout, and the run still equals the
    one-step-at-a-time run (counts equal, state within 1e-10 of each array's
    scale: the FP64 force sums are atomic)."""
    P, V = slot_lattice()
    L = np.array([12.6, 12.6, 12.6])
    say("run with a slot overflow at step 5 fp64")
    ref, cref = state_after(eng, P, V, "p3bca", 2.0, "fp64", [("step", 1)] * 9, L=L)
    eng.set_cell_slots(8)
    try:
        state_after(eng, P, V, "p3bca", 2.0, "fp64", [("run", 4)], L=L)
        assert eng.get_cell_slots() == (8, 8)  # four steps on 8 slots per cell
        got, cgot = state_after(eng, P, V, "p3bca", 2.0, "fp64", [("run", 9)], L=L)
        assert eng.get_cell_slots() == (8, 0)  # a cell outgrew its slots
        assert_same(got, cgot, ref, cref, False, tol=1e-10)
    finally:
        eng.set_cell_slots(-1)


def test_cell_slots_setting(eng):
    """Slots off, automatic or set give the same counts; values outside
    -1 .. 256 are refused."""
    P = uniform_particles(3000, (1500.0) ** (1 / 3), 44)
    V = np.zeros_like(P)
    counts = {}
    for slots in (0, -1, 64):
        eng.set_cell_slots(slots)
        _, counts[slots] = state_after(eng, P, V, "p3bca", 2.0, "fp64", [("run", 3)])
        req, use = eng.get_cell_slots()
        assert req == slots and (use == 0) == (slots == 0) and (slots <= 0 or use == slots)
    eng.set_cell_slots(-1)
    assert counts[0] == counts[-1] == counts[64]
    for bad in (-2, 257):
        with pytest.raises(tb3.TB3Error):
            eng.set_cell_slots(bad)


def lattice_cluster(n, h, centre, rng):
    """n sites of a cubic lattice of spacing h nearest the centre, jittered by
    at most h/10: a dense cluster without close pairs."""
    k = int(np.ceil((n * 6 / np.pi) ** (1 / 3))) + 2
    g = (np.arange(k) - (k - 1) / 2) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    X = X[np.argsort((X ** 2).sum(1), kind="stable")[:n]]
    return X + rng.uniform(-0.1 * h, 0.1 * h, X.shape) + centre


def test_run_with_dense_overflow_midway():
    """Two clusters of 800 (radius ~0.56, cutoff 2): every owner has 799
    neighbours, more than the cell kernel's lists hold, so all owners go to
    the dense-owner kernel.  At step 5 the clusters overlap and an owner has
    ~1596 neighbours, more than the dense-owner kernel stages: the sweep
    overflows, the replays after it halt (their sweep kernels must do nothing,
    and the dense-owner queue must not grow across them) and the host
    finishes with the per-owner kernel.  nu = 1e-12 keeps the forces below 1,
    so the drift alone brings the clusters together."""
    rng = np.random.default_rng(12)
    P = np.concatenate([lattice_cluster(800, 0.1, np.array([6.0, 10.0, 10.0]), rng),
                        lattice_cluster(800, 0.1, np.array([14.0, 10.0, 10.0]), rng)])
    V = np.zeros_like(P)
    V[:800, 0] = 600.0
    V[800:, 0] = -600.0
    L = np.array([20.0, 20.0, 20.0])
    say("run with a dense-owner overflow midway fp64")
    with tb3.Engine(0) as e1:
        ref, cref = state_after(e1, P, V, "p3bca", 2.0, "fp64", [("step", 1)] * 9, L=L, nu=1e-12)
    with tb3.Engine(0) as e2:
        got, cgot = state_after(e2, P, V, "p3bca", 2.0, "fp64", [("run", 9)], L=L, nu=1e-12)
    assert cref[0] == 2 * (800 * 799 * 798 // 6)  # each cluster: all its triplets
    assert max(cref) == 1600 * 1599 * 1598 // 6   # merged: all triplets
    assert_same(got, cgot, ref, cref, False)


def test_autotune_keeps_results(eng):
    """The tuner only changes the cell width and the persistent grid: the
    triplet set is the same and the forces agree to rounding."""
    P = uniform_particles(4000, 12.6, 5)
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=2.0, precision="fp64")
    c0 = eng.compute_forces()
    f0 = eng.get_particles()["f"]
    say("autotune p3bca n=4000")
    t = eng.autotune(reps=1)
    say(f"  tuned {t}")
    assert t["cell_ratio"] in (1.0, 1.15, 1.3, 1.5, 1.75, 2.0) and t["sweep_ms"] > 0
    c1 = eng.compute_forces()
    f1 = eng.get_particles()["f"]
    assert c1 == c0
    den = np.linalg.norm(f0, axis=1)
    num = np.linalg.norm(f1 - f0, axis=1)
    assert (num / np.maximum(den, np.sqrt(np.mean(den ** 2)))).max() <= 1e-12


def dense_core_system():
    """Uniform density-2 particles around a Gaussian core whose owners
    outgrow the cell kernel's staging (they go to the dense-owner kernel)."""
    rng = np.random.default_rng(17)
    P = np.concatenate([uniform_particles(4000, 12.6, 21), rng.normal(6.3, 0.9, (400, 3))])
    return P


@pytest.mark.parametrize("waves", [4, 8, 16])
def test_dense_waves_same_forces(eng, oracle, waves):
    """Each workgroup size of the dense-owner kernel gives the reference's
    counts and forces within the FP64 bar, from one step and from a replayed
    run (whose captured graph must pick up the new size)."""
    P = dense_core_system()
    L, _ = oracle.domain_size(P)
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=2.0, precision="fp64", domain_L=L)
    eng.set_tuning(1.0, waves)
    assert eng.get_tuning()["dense_waves"] == waves
    say(f"dense waves {waves}: p3bca n={len(P)}")
    cnt = eng.compute_forces()
    f = eng.get_particles()["f"]
    *_, fo, co = oracle.sweep(P, cutoff=2.0, L=L, mode="owner")
    assert cnt == co
    den = np.linalg.norm(fo, axis=1)
    num = np.linalg.norm(f - fo, axis=1)
    assert (num / np.maximum(den, np.sqrt(np.mean(den ** 2)))).max() <= 1e-12
    # a replayed step after a switch of the size: same counts as stepping
    V = np.zeros_like(P)
    eng.set_particles(P, V, V, np.ones(len(P)))
    c_run = eng.run(2, 0.001, (0.0, -9.81, 0.0))
    eng.set_tuning(1.0, 8 if waves != 8 else 4)
    eng.set_particles(P, V, V, np.ones(len(P)))
    c_other = eng.run(2, 0.001, (0.0, -9.81, 0.0))
    assert [int(x) for x in c_run][:1] == [int(x) for x in c_other][:1] == [co]


def test_autotune_dense_waves(eng):
    """Where the dense-owner kernel runs, the tuner also picks its waves per
    owner; the counts stay the same."""
    P = dense_core_system()
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=2.0, precision="fp64")
    c0 = eng.compute_forces()
    t = eng.autotune(reps=1)
    say(f"  tuned {t}")
    assert t["dense_waves"] in (4, 8, 16) and t["sweep_ms"] > 0
    assert eng.compute_forces() == c0
    with pytest.raises(tb3.TB3Error):
        eng.set_tuning(1.0, 12)
    with pytest.raises(tb3.TB3Error):
        eng.set_tuning(0.5, 8)


def test_run_after_new_particles(eng):
    """A new particle set in an unchanged grid (fixed domain, same count)
    reuses the captured step graph; another particle count must capture
    again.  Each run equals a fresh engine's one-step-at-a-time run, bit for
    bit on the exact path."""
    L = np.array([12.0, 12.0, 12.0])
    sets = [uniform_particles(3000, 12.0, 11), uniform_particles(3000, 12.0, 12),
            uniform_particles(2500, 12.0, 13), uniform_particles(3000, 12.0, 12)]
    eng.set_particles(sets[0], np.zeros_like(sets[0]), np.zeros_like(sets[0]), np.ones(len(sets[0])))
    eng.configure("p3bca", cutoff=2.0, precision="exact", domain_L=L)
    eng.run(4, 0.001, G)
    for k, P in enumerate(sets[1:]):
        say(f"run after set_particles #{k + 1} n={len(P)}")
        Z = np.zeros_like(P)
        eng.set_particles(P, Z, Z, np.ones(len(P)))
        counts = [int(c) for c in eng.run(4, 0.001, G)]
        st = eng.get_particles()
        with tb3.Engine(0) as fresh:
            ref, rc = state_after(fresh, P, Z, "p3bca", 2.0, "exact", [("step", 1)] * 4, L=L)
        assert_same(st, counts, ref, rc, exact=True)
