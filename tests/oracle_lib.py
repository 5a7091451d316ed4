"""ctypes binding to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests call it on the same inputs they hand to the
HIP library and compare.  Nothing in the product imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
ORACLE_CLI = os.path.join(ORACLE_DIR, "tb3_oracle_cli")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    """Compile the oracle if the shared library or CLI is missing or stale."""
    srcs = [os.path.join(ORACLE_DIR, f) for f in ("tb3_oracle.cpp", "tb3_oracle.h", "tb3_oracle_main.cpp")]
    outs = [ORACLE_SO, ORACLE_CLI]
    newest = max(os.path.getmtime(s) for s in srcs)
    if all(os.path.exists(o) and os.path.getmtime(o) >= newest for o in outs):
        return
    subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_at_contrib.argtypes = [_dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_sqr_dist_periodic.argtypes = [_dp, _dp, _dp]
        L.orc_sqr_dist_periodic.restype = ctypes.c_double
        L.orc_sweep_brute.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp,
                                      _u64p]
        L.orc_sweep_brute.restype = ctypes.c_uint64
        L.orc_sweep_owner.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp,
                                      ctypes.c_int]
        L.orc_sweep_owner.restype = ctypes.c_uint64
        L.orc_sum_roles.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp]
        L.orc_predictor.argtypes = [ctypes.c_int64, _dp, _dp, _dp, ctypes.c_double]
        L.orc_update.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_domain_size.argtypes = [ctypes.c_int64, _dp, _dp, _dp]
        L.orc_simulate.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double, _u64p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_energy.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _u64p, ctypes.c_int]
        L.orc_energy.restype = ctypes.c_double
        L.orc_forces_fast.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _u64p,
                                      ctypes.c_int]
        L.orc_forces_fast.restype = ctypes.c_double
        L.orc_energy_scale.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _u64p,
                                       ctypes.c_int]
        L.orc_energy_scale.restype = ctypes.c_double
        L.orc_force_scale.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.orc_force_scale.restype = ctypes.c_double
        L.orc_sweep_cells_fast.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _u64p,
                                           ctypes.c_int]
        L.orc_sweep_cells_fast.restype = ctypes.c_double
        L.orc_sweep_brute_rows.argtypes = [ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                           _dp, ctypes.c_int64, _u64p]
        L.orc_sweep_brute_rows.restype = ctypes.c_uint64
        L.orc_sweep_ranks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_double, _dp, _u64p,
                                      _u64p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64)]
        L.orc_sweep_ranks.restype = ctypes.c_int
        L.orc_sweep_ranks_p3bca.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int64, _dp,
                                            ctypes.c_double, ctypes.c_double, _dp, _u64p, _u64p,
                                            ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_int64)]
        L.orc_sweep_ranks_p3bca.restype = ctypes.c_int
        L.orc_simulate_ranks_p3bca.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int64, _dp, _dp,
                                               _dp, _dp, _dp, ctypes.c_int,
                                               ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double, _u64p, _u64p,
                                               ctypes.c_void_p, ctypes.c_void_p]
        L.orc_simulate_ranks_p3bca.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def at_contrib(pi, pj, pk, nu=1.0):
    """Contributions the reference subtracts from i.f, j.f, k.f for one triplet."""
    ci, cj, ck = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().orc_at_contrib(_p(_c(pi)), _p(_c(pj)), _p(_c(pk)), nu, _p(ci), _p(cj), _p(ck))
    return ci, cj, ck


def sqr_dist_periodic(a, b, L):
    return lib().orc_sqr_dist_periodic(_p(_c(a)), _p(_c(b)), _p(_c(L)))


def sweep(pos, nu=1.0, cutoff=-1.0, L=None, mode="owner", threads=0):
    """One force sweep: returns (f0, f1, f2, f, count) with f = f0 + (f1 + f2)."""
    pos = _c(pos)
    n = pos.shape[0]
    f0, f1, f2, f = (np.zeros((n, 3)) for _ in range(4))
    Lp = _p(_c(L)) if L is not None else None
    if mode == "brute":
        poss = ctypes.c_uint64(0)
        cnt = lib().orc_sweep_brute(n, _p(pos), nu, cutoff, Lp, _p(f0), _p(f1), _p(f2), ctypes.byref(poss))
    else:
        cnt = lib().orc_sweep_owner(n, _p(pos), nu, cutoff, Lp, _p(f0), _p(f1), _p(f2), threads)
    lib().orc_sum_roles(n, _p(f0), _p(f1), _p(f2), _p(f))
    return f0, f1, f2, f, int(cnt)


def _dims(ranks, dims):
    d = (ranks,) if dims is None else tuple(int(x) for x in dims)
    return len(d), (ctypes.c_int * len(d))(*d), int(np.prod(d))


def sweep_ranks(pos, algo, ranks, nu=1.0, cutoff=-1.0, dims=None):
    """The step sweep of NATA, AUTA or P3BCA (with `cutoff`, on the Cartesian
    decomposition `dims`, (ranks,) by default) as `mpiexec -n ranks` runs it,
    restated in one process: returns (f, buffer_counts, particle_counts,
    processed) with processed a (k, 4) array of (rank, owner of b0, b1, b2)
    rows."""
    if algo == "p3bca":
        nd, dv, ranks = _dims(ranks, dims)
    pos = _c(pos)
    n = pos.shape[0]
    f = np.zeros((n, 3))
    bc = np.zeros(ranks, dtype=np.uint64)
    pc = np.zeros(ranks, dtype=np.uint64)
    cap = 4 * (ranks + 2) ** 3
    proc = np.zeros((cap, 4), dtype=np.int32)
    nproc = ctypes.c_int64(0)
    if algo == "p3bca":
        rc = lib().orc_sweep_ranks_p3bca(nd, dv, n, _p(pos), nu, cutoff, _p(f), bc.ctypes.data_as(_u64p),
                                         pc.ctypes.data_as(_u64p), proc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         cap, ctypes.byref(nproc))
    else:
        rc = lib().orc_sweep_ranks(ALGO[algo], ranks, n, _p(pos), nu, _p(f), bc.ctypes.data_as(_u64p),
                                   pc.ctypes.data_as(_u64p), proc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap,
                                   ctypes.byref(nproc))
    if rc != 0:
        raise ValueError(f"the {algo} sweep at {ranks} ranks returned {rc}")
    assert nproc.value <= cap
    return f, bc, pc, proc[:nproc.value]


def simulate_ranks_p3bca(pos, vel, acc, mass, ranks, steps, dt, g, cutoff, nu=1.0, dims=None):
    """A P3BCA run on `ranks` ranks (decomposition `dims`, (ranks,) by
    default) with the particle exchange between steps: returns (pos, vel,
    acc, f, buffer_counts, particle_counts), the counts as (steps, ranks)
    arrays."""
    nd, dv, ranks = _dims(ranks, dims)
    pos, vel, acc = (_c(a).copy() for a in (pos, vel, acc))
    n = pos.shape[0]
    f = np.zeros((n, 3))
    bc = np.zeros((steps, ranks), dtype=np.uint64)
    pc = np.zeros((steps, ranks), dtype=np.uint64)
    rc = lib().orc_simulate_ranks_p3bca(nd, dv, n, _p(pos), _p(vel), _p(acc), _p(_c(mass)), _p(f), steps, dt,
                                        _p(_c(g)), cutoff, nu, bc.ctypes.data_as(_u64p), pc.ctypes.data_as(_u64p),
                                        None, None)
    if rc != 0:
        raise ValueError(f"the p3bca run at {ranks} ranks returned {rc}")
    return pos, vel, acc, f, bc, pc


def brute_possible(pos, cutoff=-1.0, L=None, nu=1.0):
    pos = _c(pos)
    n = pos.shape[0]
    f0, f1, f2 = (np.zeros((n, 3)) for _ in range(3))
    poss = ctypes.c_uint64(0)
    Lp = _p(_c(L)) if L is not None else None
    act = lib().orc_sweep_brute(n, _p(pos), nu, cutoff, Lp, _p(f0), _p(f1), _p(f2), ctypes.byref(poss))
    return int(act), int(poss.value)


def domain_size(pos):
    pos = _c(pos)
    L, lo = np.zeros(3), np.zeros(3)
    lib().orc_domain_size(pos.shape[0], _p(pos), _p(L), _p(lo))
    return L, lo


def predictor(pos, vel, acc, dt):
    pos = _c(pos).copy()
    lib().orc_predictor(pos.shape[0], _p(pos), _p(_c(vel)), _p(_c(acc)), dt)
    return pos


def update(f, mass, vel, acc, dt, g):
    vel = _c(vel).copy()
    acc = _c(acc).copy()
    lib().orc_update(vel.shape[0], _p(_c(f)), _p(_c(mass)), _p(vel), _p(acc), dt, _p(_c(g)))
    return vel, acc


ALGO = {"nata": 0, "p3bca": 1, "auta": 2}


def simulate(pos, vel, acc, mass, algo, steps, dt, g, cutoff=0.0, nu=1.0, mode="owner", threads=0):
    """Simulation::Start at one process; returns (pos, vel, acc, f, counts)."""
    pos, vel, acc = _c(pos).copy(), _c(vel).copy(), _c(acc).copy()
    mass = _c(mass)
    n = pos.shape[0]
    f = np.zeros((n, 3))
    counts = np.zeros(max(steps, 1), dtype=np.uint64)
    lib().orc_simulate(n, _p(pos), _p(vel), _p(acc), _p(mass), _p(f), ALGO[algo] if isinstance(algo, str) else algo,
                       steps, dt, _p(_c(g)), cutoff, nu, counts.ctypes.data_as(_u64p), 0 if mode == "brute" else 1,
                       threads, None, None)
    return pos, vel, acc, f, counts[:steps]


def energy(pos, nu=1.0, cutoff=-1.0, L=None, threads=0):
    pos = _c(pos)
    cnt = ctypes.c_uint64(0)
    Lp = _p(_c(L)) if L is not None else None
    E = lib().orc_energy(pos.shape[0], _p(pos), nu, cutoff, Lp, ctypes.byref(cnt), threads)
    return E, int(cnt.value)


def forces_fast(pos, nu=1.0, cutoff=-1.0, L=None, threads=0):
    pos = _c(pos)
    f = np.zeros_like(pos)
    cnt = ctypes.c_uint64(0)
    Lp = _p(_c(L)) if L is not None else None
    E = lib().orc_forces_fast(pos.shape[0], _p(pos), nu, cutoff, Lp, _p(f), ctypes.byref(cnt), threads)
    return f, E, int(cnt.value)


def energy_scale(pos, nu=1.0, cutoff=-1.0, L=None, threads=0):
    """Sum over the triplets of |nu| (1 + |3 cos cos cos|) / r^3: the
    condition scale of the energy sum."""
    pos = _c(pos)
    cnt = ctypes.c_uint64(0)
    Lp = _p(_c(L)) if L is not None else None
    return lib().orc_energy_scale(pos.shape[0], _p(pos), nu, cutoff, Lp, ctypes.byref(cnt), threads)


def force_scale(pos, nu=1.0, cutoff=-1.0, L=None, threads=0):
    """Per particle, the summed lengths of the reference's triplet
    contributions to its force (the condition scale of its force sum)."""
    pos = _c(pos)
    s = np.zeros(pos.shape[0])
    Lp = _p(_c(L)) if L is not None else None
    lib().orc_force_scale(pos.shape[0], _p(pos), nu, cutoff, Lp, _p(s), threads)
    return s


def sweep_cells_fast(pos, nu=1.0, cutoff=2.0, L=None, threads=0):
    pos = _c(pos)
    f = np.zeros_like(pos)
    cnt = ctypes.c_uint64(0)
    E = lib().orc_sweep_cells_fast(pos.shape[0], _p(pos), nu, cutoff, _p(_c(L)), _p(f), ctypes.byref(cnt), threads)
    return f, E, int(cnt.value)
