"""GPU parity at the workload configurations' own sizes, against the oracle.

Each test runs the HIP library through the C ABI on a configuration's full
particle count and compares with the oracle's one-process run
(Simulation.cpp:28-81; the sweep of Algorithm.cpp:72-139 with P3BCA's
periodic cutoff test, structs.hpp:106-115, at p = 1: P3BCA.cpp:516-805,
AUTA.cpp:268-302) on the same inputs:

  * P3BCA c=2, N=32768 uniform in 25.4^3 (reference generator, seed
    926762934), 10 steps at dt 0.001 with g = (0, -9.81, 0): the bit-exact
    path's positions, velocities, accelerations, forces and per-step counts
    identical to the oracle's; the fast FP64 sweep of the initial state with
    exact counts, its per-particle force error stated three ways; the
    FP32-mixed sweep against the force scale (below).
  * AUTA, N=2048 uniform in 10^3, one step: bit-exact path identical; fast
    FP64 sweep within the FP64 bars.
  * AUTA FP32-mixed on the reference's 25-cluster Gaussian generator
    (centres uniform in 20^3, offsets N(0, 1)), N=1000.
  * P3BCA c=2.5 FP64 energy and adjoint on the 8000-particle jittered grid
    (20^3, spacing 0.8, N(0, 0.05)) in a 16^3 domain: energy, the exact
    adjoint bit-identical to the oracle's force, dE/dnu, and central
    differences on 64 coordinates.
  * 64 replicas of the batched configuration (N log-uniform in [256, 4096],
    one Gaussian cloud of deviation 1.5 each): FP32-mixed energies, counts and
    forces of every 8th replica against the oracle; 4 FP32-mixed steps of the
    batch against every replica run alone (step 0's counts identical, later
    ones to a few per 1e5); 2 FP64 steps of the batch against every replica
    alone (counts identical, positions to 1e-9).

Error measures of a fast path's force f against the reference's fo:
  max_rel         max_p |f_p - fo_p| / |fo_p|
  rms_rel         RMS over p of the same ratio
  max_rel_floored max_p |f_p - fo_p| / max(|fo_p|, RMS |fo|)
  max_rel_scale   max_p |f_p - fo_p| / S_p, S_p = sum over p's triplets of
                  |contribution to p| (oracle.force_scale).  A force that is
                  a sum of terms of total length S_p, each computed with
                  relative error e, is off by at most e S_p; |fo_p| can be far
                  smaller than S_p where the terms cancel, which is what the
                  pure ratio then measures.  This is the FP32-mixed bar.
"""
import sys

import numpy as np
import pytest

import threebodymdalgorithms_amd as tb3
from threebodymdalgorithms_amd import generators
from conftest import force_longdouble

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

G = np.array([0.0, -9.81, 0.0])
DT = 0.001
SEED = 926762934


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = tb3.Engine(0)
    yield e
    e.close()


def errors(f, fo, scale=None):
    num = np.linalg.norm(f - fo, axis=1)
    den = np.linalg.norm(fo, axis=1)
    rel = num / np.where(den > 0, den, 1.0)
    floor = np.sqrt(np.mean(den ** 2))
    out = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2))),
           "max_rel_floored": float((num / np.maximum(den, floor)).max())}
    if scale is not None:
        out["max_rel_scale"] = float((num / np.where(scale > 0, scale, 1.0)).max())
    return out


def c1_particles():
    return generators.csv_rounded(generators.uniform(32768, 25.4, SEED))


def test_c1_exact_ten_steps_bit_identical(eng, oracle):
    P = c1_particles()
    n = len(P)
    V, A, M = np.zeros_like(P), np.zeros_like(P), np.ones(n)
    say("C1 exact: P3BCA c=2 N=32768, 10 steps")
    eng.set_particles(P, V, A, M)
    eng.configure("p3bca", cutoff=2.0, precision="exact")
    counts = [eng.step(DT, G)[1] for _ in range(10)]
    st = eng.get_particles()
    say("  oracle ...")
    pos, vel, acc, f, oc = oracle.simulate(P, V, A, M, "p3bca", 10, DT, G, cutoff=2.0)
    say(f"  counts {counts}")
    assert counts == [int(c) for c in oc]
    for name, a, b in (("pos", st["pos"], pos), ("vel", st["vel"], vel), ("acc", st["acc"], acc),
                       ("f", st["f"], f)):
        assert np.array_equal(a, b), (name, int(np.sum(a != b)))
    # the replayed step graph gives the same trajectory
    eng.set_particles(P, V, A, M)
    eng.configure("p3bca", cutoff=2.0, precision="exact")
    rc = [int(c) for c in eng.run(10, DT, G)]
    st2 = eng.get_particles()
    assert rc == counts
    assert np.array_equal(st2["pos"], pos) and np.array_equal(st2["f"], f)


def test_c1_fast_sweeps_initial_state(eng, oracle):
    P = c1_particles()
    L, _ = oracle.domain_size(P)
    say("C1 oracle sweep and force scale")
    *_, fo, co = oracle.sweep(P, cutoff=2.0, L=L, mode="owner")
    S = oracle.force_scale(P, cutoff=2.0, L=L)
    for prec in ("fp64", "fp32"):
        eng.set_particles(P)
        eng.configure("p3bca", cutoff=2.0, precision=prec)
        cnt = eng.compute_forces()
        f = eng.get_particles()["f"]
        e = errors(f, fo, S)
        say(f"  C1 {prec}: count {cnt} (oracle {co}) errors {e}")
        assert cnt == co
        if prec == "fp64":
            # Against the reference the fast path is off by a few 1e-12 of the
            # RMS force on a few particles: the reference's own rounding.  Its
            # powers and divisions of the side lengths lose bits on triplets
            # with a close pair, whose far particle receives two huge, nearly
            # opposite terms (the squared-distance form keeps them).  An 80-bit
            # evaluation of the true force on the worst particles shows it.
            assert e["max_rel_floored"] <= 2e-11 and e["rms_rel"] <= 5e-12
            assert e["max_rel"] <= 1e-9
            num = np.linalg.norm(f - fo, axis=1)
            den = np.linalg.norm(fo, axis=1)
            floor = np.sqrt(np.mean(den ** 2))
            worst = np.argsort(num / np.maximum(den, floor))[-8:]
            rng = np.random.default_rng(3)
            for p in list(worst) + list(rng.choice(len(P), 8, replace=False)):
                truth = force_longdouble(P, int(p), 2.0, L)
                sc = max(np.linalg.norm(truth), floor)
                e_fast = np.linalg.norm(f[p] - truth) / sc
                e_ref = np.linalg.norm(fo[p] - truth) / sc
                say(f"    particle {p}: vs the 80-bit force fast {e_fast:.2e}, reference {e_ref:.2e}")
                assert e_fast <= 1e-12
        else:
            # FP32 coefficient arithmetic: 2e-7 of the RMS force, 1.5e-6 of
            # the per-particle scale on the worst particle (a triplet whose
            # gradient coefficient is a difference of larger FP32 terms)
            assert e["max_rel_scale"] <= 2e-6
            assert e["max_rel_floored"] <= 1e-6 and e["rms_rel"] <= 1e-5


def test_c1_fast_run_counts_match_exact(eng):
    """C1's replayed FP64 steps (slotted cell layout: the binning fills each
    cell's slots, no scatter pass) count the same triplets per step as the
    exact path, which is the reference bit for bit, and as replays on the
    contiguous layout."""
    P = c1_particles()
    Z = np.zeros_like(P)
    res = {}
    for prec, slots in (("exact", -1), ("fp64", -1), ("fp64", 0)):
        eng.set_cell_slots(slots)
        eng.set_particles(P, Z, Z, np.ones(len(P)))
        eng.configure("p3bca", cutoff=2.0, precision=prec)
        say(f"C1 run of 4 steps {prec}, slots {slots}")
        res[(prec, slots)] = [int(c) for c in eng.run(4, 0.001, (0.0, -9.81, 0.0))]
        assert (eng.get_cell_slots()[1] > 0) == (prec == "fp64" and slots == -1)
    eng.set_cell_slots(-1)
    say(f"  counts {res}")
    assert res[("fp64", -1)] == res[("exact", -1)] == res[("fp64", 0)]


def test_c0_auta_2048_one_step(eng, oracle):
    P = generators.csv_rounded(generators.uniform(2048, 10.0, SEED))
    n = len(P)
    V, A, M = np.zeros_like(P), np.zeros_like(P), np.ones(n)
    say("C0 exact: AUTA N=2048, one step")
    eng.set_particles(P, V, A, M)
    eng.configure("auta", precision="exact")
    nb, nt = eng.step(DT, G)
    st = eng.get_particles()
    say("  oracle ...")
    pos, vel, acc, f, oc = oracle.simulate(P, V, A, M, "auta", 1, DT, G)
    assert (nb, nt) == (1, int(oc[0])) and nt == n * (n - 1) * (n - 2) // 6
    for name, a, b in (("pos", st["pos"], pos), ("vel", st["vel"], vel), ("acc", st["acc"], acc),
                       ("f", st["f"], f)):
        assert np.array_equal(a, b), (name, int(np.sum(a != b)))
    # the fast FP64 sweep of the predicted positions (a zero-velocity,
    # zero-acceleration predictor leaves them unchanged) against that force
    eng.set_particles(P)
    eng.configure("auta", precision="fp64")
    cnt = eng.compute_forces()
    e = errors(eng.get_particles()["f"], f)
    say(f"  C0 fp64 errors {e}")
    assert cnt == nt
    # Without a cutoff a particle sums ~2e6 triplets, many with a close pair
    # far away whose two large terms nearly cancel on it; both FP64
    # formulations lose bits there.  Against an 80-bit evaluation the
    # squared-distance form is off by up to ~1e-11 of the RMS force on the
    # worst particles, the reference's expression by less than 1e-12.
    assert e["max_rel_floored"] <= 2e-11 and e["rms_rel"] <= 1e-12 and e["max_rel"] <= 2e-11


def test_c2_clustered_auta_fp32(eng, oracle):
    P = generators.csv_rounded(generators.clustered_gauss(1000, 25, 20.0, 0.0, 1.0, SEED, 89347587))
    assert len(P) == 1000
    say("C2 clustered AUTA N=1000 FP32-mixed: oracle ...")
    *_, fo, co = oracle.sweep(P, mode="owner")
    S = oracle.force_scale(P)
    Eo, _ = oracle.energy(P)
    Es = oracle.energy_scale(P)
    eng.set_particles(P)
    eng.configure("auta", precision="fp32")
    E, g, dnu, cnt = eng.energy_grad()
    e = errors(g, fo, S)
    say(f"  C2 fp32: count {cnt} energy error {abs(E - Eo) / abs(Eo):.2e} of |E|, {abs(E - Eo) / Es:.2e} of the "
        f"energy scale; force errors {e}")
    assert cnt == co
    assert e["max_rel_scale"] <= 2e-6
    assert e["max_rel_floored"] <= 1e-6 and e["rms_rel"] <= 1e-5
    assert abs(E - Eo) <= 1e-6 * Es
    eng.configure("auta", precision="fp64")
    E64, g64, _, _ = eng.energy_grad()
    e64 = errors(g64, fo, S)
    say(f"  C2 fp64 errors {e64}; energy {abs(E64 - Eo) / Es:.2e} of the energy scale")
    # FP64 against the reference: the usual per-sweep bars; the energy is a
    # sum of 1.7e8 terms in two formulations and two summation orders
    assert e64["max_rel_floored"] <= 1e-12 and e64["rms_rel"] <= 1e-12 and e64["max_rel"] <= 1e-10
    assert e64["max_rel_scale"] <= 1e-11 and abs(E64 - Eo) <= 1e-11 * Es


def jittered_grid(seed):
    g = np.arange(20) * 0.8
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    return P + np.random.default_rng(seed).normal(0.0, 0.05, P.shape)


def test_c3_energy_adjoint_8000(eng, oracle):
    torch = pytest.importorskip("torch")
    from threebodymdalgorithms_amd.autograd import at_energy
    P = jittered_grid(0)
    L = np.full(3, 16.0)
    say("C3 P3BCA c=2.5 N=8000 energy + adjoint")
    Eo, co = oracle.energy(P, cutoff=2.5, L=L)
    *_, fo, _ = oracle.sweep(P, cutoff=2.5, L=L, mode="owner")
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=2.5, domain_L=L, precision="exact")
    Ee, ge, dnu_e, ce = eng.energy_grad()
    assert ce == co
    assert np.array_equal(ge, fo)
    assert Ee == pytest.approx(Eo, rel=1e-12) and dnu_e == pytest.approx(Ee, rel=1e-15)

    eng.configure("p3bca", cutoff=2.5, domain_L=L, precision="fp64")
    x = torch.tensor(P, dtype=torch.float64, device="cuda", requires_grad=True)
    nu = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
    E = at_energy(x, nu, eng)
    E.backward()
    g = x.grad.detach().cpu().numpy()
    e = errors(g, fo)
    say(f"  C3 energy {float(E.detach()):.10e} (oracle {Eo:.10e}); fp64 adjoint errors {e}")
    assert float(E.detach()) == pytest.approx(Eo, rel=1e-12)
    assert float(nu.grad) == pytest.approx(float(E.detach()), rel=1e-14)
    assert e["max_rel_floored"] <= 1e-12 and e["rms_rel"] <= 1e-12

    # central differences (h = 1e-6) on 64 coordinates whose triplet set is
    # unchanged over the stencil.  Moving particle i changes only the
    # triplets that contain i, and those lie within the cutoff of i, so the
    # difference of the full energy equals the difference of the energy of
    # i and its neighbours (same periodic domain): the difference quotient is
    # taken of that small sum, of the GPU's FP64 energy and of the oracle's
    # textbook energy, free of the rounding of the full 3e5-sized sum.
    rng = np.random.default_rng(64)
    h = 1e-6
    gmax = np.abs(g).max()
    checked = 0
    worst = 0.0
    for _ in range(400):
        i, c = int(rng.integers(len(P))), int(rng.integers(3))
        dd = np.abs(P - P[i])
        dd = np.minimum(dd, L - dd)
        near = np.flatnonzero(np.sum(dd * dd, axis=1) <= 2.6 ** 2)
        sub = P[near]
        me = int(np.flatnonzero(near == i)[0])
        Eg, Eo_, cnt = [], [], []
        for d in (h, -h):
            Q = sub.copy()
            Q[me, c] += d
            eng.set_particles(Q)
            Ev, cv = eng.energy()
            Ex, cx = oracle.energy(Q, cutoff=2.5, L=L)
            Eg.append(Ev)
            Eo_.append(Ex)
            cnt.append((cv, cx))
        eng.set_particles(sub)
        _, c_mid = eng.energy()
        if any(cv != c_mid or cx != c_mid for cv, cx in cnt):
            continue
        fd_gpu = (Eg[0] - Eg[1]) / (2 * h)
        fd_orc = (Eo_[0] - Eo_[1]) / (2 * h)
        tol = 1e-6 * max(abs(g[i, c]), 0.01 * gmax)
        assert abs(fd_gpu - g[i, c]) <= tol, (i, c, fd_gpu, g[i, c])
        assert abs(fd_orc - g[i, c]) <= tol, (i, c, fd_orc, g[i, c])
        worst = max(worst, abs(fd_gpu - g[i, c]) / max(abs(g[i, c]), 0.01 * gmax))
        checked += 1
        if checked == 64:
            break
    say(f"  C3 central differences on 64 coordinates: worst relative deviation {worst:.2e}")
    assert checked == 64


def test_c4_batch_64_replicas(eng, oracle):
    rng = np.random.default_rng(SEED)
    sizes = np.rint(np.exp(rng.uniform(np.log(256), np.log(4096), 64))).astype(int)
    reps = [generators.csv_rounded(generators.gauss(int(n), 0.0, 1.5, 1000 + r)) for r, n in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    P = np.concatenate(reps)
    say(f"C4 batch: 64 replicas, {len(P)} particles")
    Z = np.zeros_like(P)
    eng.set_particles(P, Z, Z, np.ones(len(P)))
    eng.configure("p3bca", cutoff=2.0, precision="fp32")
    eng.batch_set(off)
    # FP32-mixed energies, counts and forces of the initial state, replica by
    # replica against the oracle (each replica on its own bbox-rule domain)
    E0, c0 = eng.batch_energy()
    eng.compute_forces()
    f32 = eng.get_particles()["f"]
    worst_e = worst_f = worst_fl = 0.0
    for r in range(0, 64, 8):
        Q = reps[r]
        L, _ = oracle.domain_size(Q)
        Eo, co = oracle.energy(Q, cutoff=2.0, L=L)
        Es = oracle.energy_scale(Q, cutoff=2.0, L=L)
        *_, fo, _ = oracle.sweep(Q, cutoff=2.0, L=L, mode="owner")
        S = oracle.force_scale(Q, cutoff=2.0, L=L)
        assert int(c0[r]) == co, r
        # each FP32 term carries ~1e-7 relative error in each of its two parts
        assert abs(E0[r] - Eo) <= 1e-6 * Es, (r, E0[r], Eo, Es)
        e = errors(f32[off[r]:off[r + 1]], fo, S)
        assert e["max_rel_scale"] <= 2e-6, (r, e)
        worst_e = max(worst_e, abs(E0[r] - Eo) / Es)
        worst_f, worst_fl = max(worst_f, e["max_rel_scale"]), max(worst_fl, e["max_rel_floored"])
    say(f"  C4 fp32 vs oracle (8 replicas): energy {worst_e:.2e} of the energy scale, force {worst_f:.2e} of "
        f"the force scale, {worst_fl:.2e} floored at each replica's RMS force")

    # four FP32-mixed steps of the batch against each replica run alone: the
    # step-0 sweep counts the same triplets; FP32 rounding turns the last-bit
    # FP64 differences of the two summation orders into FP32-ulp jumps, and
    # from step 1 on a pair within ~1e-10 of the cutoff (the densest replicas
    # hold ~1e7 triplets) can count differently, so later counts agree to a
    # few per 1e5
    eng.set_particles(P, Z, Z, np.ones(len(P)))
    eng.configure("p3bca", cutoff=2.0, precision="fp32")
    eng.batch_set(off)
    bc = eng.batch_run(4, DT, G)
    count_dev = 0.0
    for r in range(64):
        Q = reps[r]
        eng.set_particles(Q, np.zeros_like(Q), np.zeros_like(Q), np.ones(len(Q)))
        eng.configure("p3bca", cutoff=2.0, precision="fp32")
        counts = [eng.step(DT, G)[1] for _ in range(4)]
        assert int(bc[0, r]) == counts[0], r
        for k in (1, 2, 3):
            assert abs(int(bc[k, r]) - counts[k]) <= 1e-3 * counts[k], (r, k)
            count_dev = max(count_dev, abs(int(bc[k, r]) - counts[k]) / max(counts[k], 1))
    say(f"  C4 fp32 per-step totals {bc.sum(axis=1).tolist()}; batch vs alone counts of steps 1-3 within "
        f"{count_dev:.1e}")

    # two FP64 steps of the batch against each replica alone: forces agree to
    # ~1e-13 of the RMS force, so positions stay within ~1e-10 even for
    # accelerations of 1e9 (bar 1e-9 of the RMS position) and counts agree
    eng.set_particles(P, Z, Z, np.ones(len(P)))
    eng.configure("p3bca", cutoff=2.0, precision="fp64")
    eng.batch_set(off)
    bc = eng.batch_run(2, DT, G)
    st = eng.get_particles()
    worst_x = 0.0
    for r in range(64):
        Q = reps[r]
        eng.set_particles(Q, np.zeros_like(Q), np.zeros_like(Q), np.ones(len(Q)))
        eng.configure("p3bca", cutoff=2.0, precision="fp64")
        counts = [eng.step(DT, G)[1] for _ in range(2)]
        alone = eng.get_particles()
        assert [int(c) for c in bc[:, r]] == counts, r
        ex = float(np.max(np.abs(st["pos"][off[r]:off[r + 1]] - alone["pos"])) / np.sqrt(np.mean(alone["pos"] ** 2)))
        worst_x = max(worst_x, ex)
        assert ex <= 1e-9, (r, ex)
    say(f"  C4 fp64 batch vs alone after 2 steps: positions within {worst_x:.2e} of the RMS position")
