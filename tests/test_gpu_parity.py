"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on
the same seeded inputs.

Bars (written into each test):
  * TB3_FP64_EXACT: bit-identical forces, positions, velocities, accelerations
    and triplet counts, over every step.
  * TB3_FP64 (reformulated, reordered): one sweep from identical positions,
    per-particle |df| <= 1e-12 max(|f|, rms|f|) and RMS of |df|/|f| <= 1e-12;
    counts bit-exact.  (A particle whose net force nearly cancels amplifies
    ulp-level differences in the pure ratio |df|/|f|, which is held to 1e-10.)
  * TB3_FP32_MIXED: one sweep, |df| <= 2e-6 of the particle's summed
    contribution lengths, |df| <= 1e-6 max(|f|, rms|f|), RMS |df|/|f| <= 1e-5;
    counts bit-exact (the cutoff test is FP64).  bench.py states the measured
    errors.
"""
import sys

import numpy as np
import pytest

import threebodymdalgorithms_amd as tb3
from conftest import clustered_particles, uniform_particles

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = tb3.Engine(0)
    yield e
    e.close()


def per_particle_rel(a, b):
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return np.max(num / np.where(den > 0, den, 1.0))


def force_errors(f, fo, scale=None):
    """Per-particle force errors of f against the reference fo: pure ratio
    |df|/|F|, its RMS, the ratio with |F| floored at the RMS force, and, given
    each particle's summed contribution lengths S_p (oracle.force_scale),
    max |df|/S_p: a sum of terms of total length S_p computed with relative
    error e per term is off by at most e S_p, however much the terms cancel."""
    num = np.linalg.norm(f - fo, axis=1)
    den = np.linalg.norm(fo, axis=1)
    rel = num / np.where(den > 0, den, 1.0)
    floor = np.sqrt(np.mean(den ** 2))
    out = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel ** 2))),
           "max_rel_floored": float((num / np.maximum(den, floor)).max())}
    if scale is not None:
        out["max_rel_scale"] = float((num / np.where(scale > 0, scale, 1.0)).max())
    return out


def gpu_forces(eng, P, algo, cutoff, precision, nu=1.0, L=None):
    eng.set_particles(P)
    eng.configure(algo, nu=nu, cutoff=cutoff, domain_L=L, precision=precision)
    cnt = eng.compute_forces()
    return eng.get_particles()["f"], cnt


def test_fp64_sqrt_and_division_round_correctly(eng):
    """The exact path needs IEEE-rounded FP64 sqrt and division on the device."""
    rng = np.random.default_rng(0)
    n = 2_000_000
    a = np.concatenate([rng.uniform(0, 1, n // 4), 10.0 ** rng.uniform(-30, 30, n // 4),
                        rng.uniform(1, 4, n // 4), np.abs(rng.normal(size=n // 4)) * 1e8])
    b = np.concatenate([rng.uniform(0.5, 2, n // 2), 10.0 ** rng.uniform(-20, 20, n // 2)]) * \
        rng.choice([-1.0, 1.0], n)
    say(f"selftest_fp64 n={n}")
    sq, q, rs = eng.selftest_fp64(a, b)
    assert np.array_equal(sq, np.sqrt(a)), int(np.sum(sq != np.sqrt(a)))
    assert np.array_equal(q, a / b), int(np.sum(q != a / b))
    # the hardware 1/sqrt estimate seeds the fast path's Newton steps
    rel = np.abs(rs * np.sqrt(a) - 1.0)
    say(f"  v_rsq_f64 max relative error {rel.max():.3e} (2^{np.log2(max(rel.max(), 1e-300)):.1f})")
    assert rel.max() < 1e-6


def test_fp64_inrange_sequences_match_ieee(eng):
    """The exact path's division and square root without the range-scaling
    steps give the IEEE results' bits on operands within 2^-500 .. 2^500: the
    numerators (small integers, 5 d, 3 nu) and denominators (products of up to
    the 11th power of a side) the brackets form, random mantissas with
    exponents across that range, values next to powers of two, and all-ones
    mantissas."""
    rng = np.random.default_rng(1)
    n = 1_000_000
    # a and b within 2^-250 .. 2^250, so that a / b is within 2^-500 .. 2^500
    a = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-250, 250, n)) * rng.choice([-1.0, 1.0], n)
    b = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-250, 250, n)) * rng.choice([-1.0, 1.0], n)
    d = 2.0 ** rng.uniform(-40, 40, n)
    small = rng.choice([-8.0, 1.0, 3.0, 5.0, 6.0], n)
    pw = rng.integers(1, 12, n)
    a2 = np.where(rng.random(n) < 0.5, small, 5.0 * d)
    b2 = (d ** pw) * 2.0 ** rng.uniform(-1, 1, n)
    e = rng.integers(-245, 245, 200_000).astype(float)
    edge = np.concatenate([2.0 ** e, np.nextafter(2.0 ** e, 0), np.nextafter(2.0 ** e, np.inf),
                           (2.0 - 2.0 ** -52) * 2.0 ** e])
    A = np.concatenate([a, a2, edge, rng.permutation(edge)])
    B = np.concatenate([b, b2, rng.permutation(edge), edge])
    say(f"in-range sequences n={A.size}")
    # square roots over 2^-500 .. 2^500
    S = np.concatenate([np.abs(A), np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-500, 500, n))])
    sq, _ = eng.selftest_fp64_inrange(S, np.ones_like(S))
    ref_sq, _, _ = eng.selftest_fp64(S, np.ones_like(S))
    assert np.array_equal(sq, ref_sq), int(np.sum(sq != ref_sq))
    assert np.array_equal(sq, np.sqrt(S))
    _, q = eng.selftest_fp64_inrange(A, B)
    _, ref_q, _ = eng.selftest_fp64(A, B)
    assert np.array_equal(q, ref_q), int(np.sum(q != ref_q))
    assert np.array_equal(q, A / B)


@pytest.mark.parametrize("algo,n,box,cutoff", [("auta", 150, 4.0, 0.0), ("nata", 97, 4.0, 0.0),
                                               ("p3bca", 800, 9.0, 2.0), ("p3bca", 300, 5.0, 0.0),
                                               ("p3bca", 400, 4.5, 2.4)])
def test_exact_forces_bit_identical(eng, oracle, algo, n, box, cutoff):
    P = uniform_particles(n, box, 100 + n)
    L, _ = oracle.domain_size(P)
    use_cut = cutoff if algo == "p3bca" else -1.0
    say(f"exact {algo} n={n} cutoff={cutoff}")
    f, cnt = gpu_forces(eng, P, algo, cutoff, "exact")
    *_, fo, co = oracle.sweep(P, cutoff=use_cut, L=L, mode="owner")
    assert cnt == co
    assert np.array_equal(f, fo), per_particle_rel(f, fo)


@pytest.mark.parametrize("n,box,cutoff", [(900, 5.0, 2.2), (3100, 6.7, 2.2)])
def test_exact_long_neighbour_lists(eng, oracle, n, box, cutoff):
    """Neighbour lists past 256 entries (the exact sweep sorts those with a
    bitonic sort instead of ranks) and past the initial list capacity (the
    sweep overflows once and is re-run with a larger capacity): still bit for
    bit the reference's forces and counts."""
    P = uniform_particles(n, box, 31 + n)
    L, _ = oracle.domain_size(P)
    d = np.abs(P[:, None, :] - P[None, :, :])
    d = np.minimum(d, L - d)
    longest = int(((d ** 2).sum(-1) <= cutoff ** 2).sum(1).max()) - 1
    say(f"exact p3bca n={n} box={box} cutoff={cutoff}: longest list {longest}")
    assert longest > 256
    f, cnt = gpu_forces(eng, P, "p3bca", cutoff, "exact", L=L)
    *_, fo, co = oracle.sweep(P, cutoff=cutoff, L=L, mode="owner")
    assert cnt == co
    assert np.array_equal(f, fo), per_particle_rel(f, fo)


def test_exact_forces_particles_outside_box(eng, oracle):
    """Particles displaced by one or two box lengths (the one-image rule)."""
    n = 500
    P = uniform_particles(n, 8.0, 5)
    L, _ = oracle.domain_size(P)
    rng = np.random.default_rng(6)
    idx = rng.choice(n, 60, replace=False)
    P2 = P.copy()
    P2[idx, rng.integers(0, 3, 60)] += L[0] * rng.choice([-2.0, -1.0, 1.0, 2.0], 60)
    say("exact p3bca outside-box")
    f, cnt = gpu_forces(eng, P2, "p3bca", 2.0, "exact", L=L)
    *_, fo, co = oracle.sweep(P2, cutoff=2.0, L=L, mode="owner")
    assert cnt == co
    assert np.array_equal(f, fo)


def test_exact_trajectory_bit_identical(eng, oracle):
    """10 steps of P3BCA (c = 2, density 2): every state array and every count."""
    n = 3000
    box = (n / 2.0) ** (1 / 3)
    P = uniform_particles(n, box, 77)
    V, A, M = np.zeros_like(P), np.zeros_like(P), np.ones(n)
    g = np.array([0.0, -9.81, 0.0])
    eng.set_particles(P, V, A, M)
    eng.configure("p3bca", cutoff=2.0, precision="exact")
    say("exact trajectory p3bca n=3000 x10")
    counts = [eng.step(0.001, g)[1] for _ in range(10)]
    st = eng.get_particles()
    pos, vel, acc, f, oc = oracle.simulate(P, V, A, M, "p3bca", 10, 0.001, g, cutoff=2.0)
    assert counts == [int(c) for c in oc]
    for name, a, b in (("pos", st["pos"], pos), ("vel", st["vel"], vel), ("acc", st["acc"], acc),
                       ("f", st["f"], f)):
        assert np.array_equal(a, b), name


def test_large_grid_three_kernel_cell_list(eng, oracle):
    """Grids of more than 8192 cells build the cell list with a separate
    count scan (smaller grids scan inside the scatter kernel): cutoff 0.7 in
    a 21.5^3 box gives 30^3 cells.  Three exact steps bit for bit (the cell
    counters must come back clear after every step), then one fast FP64
    sweep with exact counts."""
    n = 20000
    box = (n / 2.0) ** (1 / 3)
    P = uniform_particles(n, box, 91)
    V, A, M = np.zeros_like(P), np.zeros_like(P), np.ones(n)
    g = np.array([0.0, -9.81, 0.0])
    eng.set_particles(P, V, A, M)
    eng.configure("p3bca", cutoff=0.7, precision="exact")
    say("exact trajectory p3bca n=20000 cutoff 0.7 (27000 cells) x3")
    counts = [eng.step(0.001, g)[1] for _ in range(3)]
    st = eng.get_particles()
    pos, vel, acc, f, oc = oracle.simulate(P, V, A, M, "p3bca", 3, 0.001, g, cutoff=0.7)
    assert counts == [int(c) for c in oc]
    for name, a, b in (("pos", st["pos"], pos), ("vel", st["vel"], vel), ("acc", st["acc"], acc),
                       ("f", st["f"], f)):
        assert np.array_equal(a, b), name
    L, _ = oracle.domain_size(P)
    say("fp64 p3bca n=20000 cutoff 0.7")
    f, cnt = gpu_forces(eng, P, "p3bca", 0.7, "fp64")
    *_, fo, co = oracle.sweep(P, cutoff=0.7, L=L, mode="owner")
    assert cnt == co
    e = force_errors(f, fo)
    say(f"  fp64 errors {e}")
    assert e["max_rel_floored"] <= 1e-12 and e["rms_rel"] <= 1e-12
    assert e["max_rel"] <= 1e-10


# 12.6: 6 cells per dimension (the cutoff tests keep the periodic form);
# 16.0: 7 cells, where the plain-difference form serves the inner cells
@pytest.mark.parametrize("algo,n,box,cutoff", [("auta", 600, 5.0, 0.0), ("p3bca", 4000, 12.6, 2.0),
                                               ("p3bca", 8192, 16.0, 2.0)])
def test_fast_fp64_one_sweep(eng, oracle, algo, n, box, cutoff):
    P = uniform_particles(n, box, 300 + n)
    L, _ = oracle.domain_size(P)
    say(f"fp64 {algo} n={n}")
    f, cnt = gpu_forces(eng, P, algo, cutoff, "fp64")
    *_, fo, co = oracle.sweep(P, cutoff=cutoff if algo == "p3bca" else -1.0, L=L, mode="owner")
    assert cnt == co
    e = force_errors(f, fo)
    say(f"  fp64 errors {e}")
    assert e["max_rel_floored"] <= 1e-12 and e["rms_rel"] <= 1e-12
    assert e["max_rel"] <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fast_sweep_particles_outside_box(eng, oracle, precision):
    """A fixed 16^3 domain (8 cells per dimension) with a fifth of the
    particles moved out of the box by +-L in one dimension: they stay in
    their wrapped cells, the cutoff test is periodic and the force uses the
    raw positions (Algorithm.cpp, AxilrodTeller.cpp).  Workgroups that
    stage such a particle must keep the periodic cutoff form; the counts
    must still match the oracle exactly."""
    rng = np.random.default_rng(77)
    L = np.array([16.0, 16.0, 16.0])
    P = uniform_particles(8192, 16.0, 7001)
    moved = rng.random(len(P)) < 0.2
    dim = rng.integers(0, 3, len(P))
    sign = rng.choice([-1.0, 1.0], len(P))
    P[moved, dim[moved]] += sign[moved] * 16.0
    say(f"{precision} p3bca n=8192 with {int(moved.sum())} particles outside the box")
    f, cnt = gpu_forces(eng, P, "p3bca", 2.0, precision, L=L)
    *_, fo, co = oracle.sweep(P, cutoff=2.0, L=L, mode="owner")
    assert cnt == co
    e = force_errors(f, fo, oracle.force_scale(P, cutoff=2.0, L=L))
    say(f"  errors {e}")
    # the moved particles' forces come from raw vectors about a box length
    # long and largely cancel, which inflates the pure ratio |df|/|F| of a
    # few particles; the ratio floored at the RMS force carries the FP64 bar.
    # (The FP64 ratio to the summed contribution lengths is only printed: a
    # triplet with a close pair hands its third particle two large, nearly
    # opposite terms whose sum the reference's expression rounds to ~1e-10 of
    # itself, so that ratio measures the reference's rounding there.)
    if precision == "fp64":
        assert e["max_rel_floored"] <= 1e-12 and e["rms_rel"] <= 1e-10
    else:
        assert e["max_rel_floored"] <= 1e-6 and e["rms_rel"] <= 1e-3
        assert e["max_rel_scale"] <= 2e-6


@pytest.mark.parametrize("algo,n,box,cutoff", [("auta", 600, 5.0, 0.0), ("p3bca", 4000, 12.6, 2.0),
                                               ("p3bca", 8192, 16.0, 2.0)])
def test_fp32_mixed_one_sweep(eng, oracle, algo, n, box, cutoff):
    P = uniform_particles(n, box, 500 + n)
    L, _ = oracle.domain_size(P)
    say(f"fp32 {algo} n={n}")
    f, cnt = gpu_forces(eng, P, algo, cutoff, "fp32")
    *_, fo, co = oracle.sweep(P, cutoff=cutoff if algo == "p3bca" else -1.0, L=L, mode="owner")
    assert cnt == co
    scale = oracle.force_scale(P, cutoff=cutoff if algo == "p3bca" else -1.0, L=L)
    e = force_errors(f, fo, scale)
    say(f"  fp32 errors {e}")
    # FP32 rounding of u, v, w (~6e-8 each) bounds each triplet's contribution
    # to ~1e-7 of its length, a few 1e-7 where a gradient coefficient is the
    # difference of larger FP32 terms (close pairs included: their third
    # particle's coefficient sum is formed without cancellation).  So the
    # error is held against each particle's summed contribution lengths; the
    # pure ratio |df|/|F| is unbounded where the contributions cancel and is
    # only printed.
    assert e["max_rel_scale"] <= 2e-6
    assert e["max_rel_floored"] <= 1e-6
    assert e["rms_rel"] <= 1e-5


@pytest.mark.parametrize("n,box,cutoff,moved", [(4000, 12.6, 2.0, False), (3000, 12.0, 2.0, True),
                                                (300, 4.0, 2.2, False)])
def test_force_scale_matches_oracle(eng, oracle, n, box, cutoff, moved):
    """tb3_force_scale (the squared-distance form, FP64) against the oracle's
    sum of the reference's contribution lengths; both bound |f| from above.
    Also particles outside the box and a grid with fewer than 3 cells."""
    P = uniform_particles(n, box, 900 + n)
    L, _ = oracle.domain_size(P)
    if moved:
        P = P.copy()
        P[::7, 0] += L[0]
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=cutoff, domain_L=L, precision="fp64")
    say(f"force scale n={n} box={box} moved={moved}")
    s = eng.force_scale()
    so = oracle.force_scale(P, cutoff=cutoff, L=L)
    *_, fo, _ = oracle.sweep(P, cutoff=cutoff, L=L, mode="owner")
    rel = np.abs(s - so) / np.where(so > 0, so, 1.0)
    say(f"  max relative difference {rel.max():.2e}")
    assert rel.max() <= 1e-9
    assert np.all(s >= np.linalg.norm(fo, axis=1) * (1 - 1e-9))
    assert np.array_equal(s == 0, so == 0)
    eng.configure("auta")
    with pytest.raises(tb3.TB3Error):
        eng.force_scale()


@pytest.mark.parametrize("precision", ["exact", "fp64", "fp32"])
def test_energy_and_adjoint(eng, oracle, precision):
    n = 1500
    P = clustered_particles(n, 6, 10.0, 1.0, 9)
    L, _ = oracle.domain_size(P)
    eng.set_particles(P)
    eng.configure("p3bca", nu=1.3, cutoff=2.5, precision=precision)
    say(f"energy {precision}")
    E, cnt = eng.energy()
    E2, g, dnu, cnt2 = eng.energy_grad()
    Eo, co = oracle.energy(P, nu=1.3, cutoff=2.5, L=L)
    *_, fo, _ = oracle.sweep(P, nu=1.3, cutoff=2.5, L=L, mode="owner")
    assert cnt == cnt2 == co
    tol = 1e-6 if precision == "fp32" else 1e-12
    assert E == pytest.approx(Eo, rel=tol)
    assert E2 == pytest.approx(Eo, rel=tol)
    assert dnu == pytest.approx(E2 / 1.3, rel=1e-15)
    if precision == "exact":
        assert np.array_equal(g, fo)
    else:
        e = force_errors(g, fo)
        say(f"  {precision} adjoint errors {e}")
        if precision == "fp32":
            assert e["max_rel_floored"] <= 1e-6 and e["rms_rel"] <= 1e-5
        else:
            assert e["max_rel_floored"] <= 1e-12 and e["rms_rel"] <= 1e-12
            assert e["max_rel"] <= 1e-10


def test_empty_and_tiny_systems(eng):
    for n in (0, 1, 2):
        P = np.arange(3 * n, dtype=float).reshape(n, 3)
        eng.set_particles(P)
        eng.configure("p3bca", cutoff=2.0, domain_L=[10, 10, 10], precision="exact")
        nb, nt = eng.step(0.001, [0, -9.81, 0])
        assert nt == 0
        st = eng.get_particles()
        assert np.array_equal(st["f"], np.zeros((n, 3)))


@pytest.mark.parametrize("precision", ["exact", "fp64"])
def test_nonfinite_position(eng, oracle, precision):
    """A NaN coordinate fails every cutoff test's "greater than", so the
    reference pairs that particle with everything (Algorithm.cpp:95-97): its
    triplets with each pair within the cutoff count, and every force they
    reach is NaN.  The cell kernel hands such a sweep to the per-owner kernel
    over all particles; counts equal the oracle's brute sweep, NaN rows
    coincide, finite rows meet the path's bar.  A replayed run (slotted cell
    layout on the fast path) that meets the NaN at its first binning counts
    the same per step as steps taken one at a time."""
    P = uniform_particles(300, 8.0, 77)
    P[17, 1] = np.nan
    L = np.array([8.0, 8.0, 8.0])
    say(f"non-finite position {precision}")
    f, cnt = gpu_forces(eng, P, "p3bca", 2.0, precision, L=L)
    *_, fo, co = oracle.sweep(P, cutoff=2.0, L=L, mode="brute")
    assert cnt == co
    bad = np.isnan(fo).any(axis=1)
    assert bad.any()  # here every particle shares a triplet with the NaN one
    assert np.array_equal(np.isnan(f).any(axis=1), bad)
    if (~bad).any():
        if precision == "exact":
            assert np.array_equal(f[~bad], fo[~bad])
        else:
            assert force_errors(f[~bad], fo[~bad])["max_rel_floored"] <= 1e-12
    V = np.zeros_like(P)
    G = np.array([0.0, -9.81, 0.0])
    eng.set_particles(P, V, V, np.ones(len(P)))
    eng.configure("p3bca", cutoff=2.0, domain_L=L, precision=precision)
    c_steps = [eng.step(0.001, G)[1] for _ in range(3)]
    eng.set_particles(P, V, V, np.ones(len(P)))
    c_run = [int(c) for c in eng.run(3, 0.001, G)]
    say(f"  counts steps {c_steps} run {c_run}")
    assert c_run == c_steps
