"""Python host interface over the C ABI.

`Engine` holds one tb3 context (one GPU, one particle system).  Its methods
mirror the reference's one-process call surface:

    reference (src/)                                  Engine
    ----------------------------------------------    -------------------------
    createNATA/P3BCA/AUTAContext (main.cpp:32-105)     configure(algo, nu, cutoff)
    getParticlesFromCSV (utility/utility.cpp:14-24)    set_particles(...) / read_csv
    Algorithm::SimulationStep (Algorithm.hpp:60)       compute_forces() -> count
    Simulation::Start iteration (Simulation.cpp:38-48) step(dt, g) -> (buffers, triplets)
    Simulation::Start loop (Simulation.cpp:30)         run(steps, dt, g) -> counts
    writeStepToCSVWithForces (utility.cpp:54-68)       get_particles() / write_csv

Errors raise TB3Error (the reference exits the process instead).
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import TB3Error


def _as(a, shape, dtype=np.float64):
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.shape != shape:
        a = a.reshape(shape)
    return a


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def device_count():
    n = ctypes.c_int(0)
    _abi.load().tb3_device_count(ctypes.byref(n))
    return n.value


class Engine:
    """One three-body force engine context on one GPU."""

    def __init__(self, device=0):
        lib = _abi.load()
        self._lib = lib
        ctx = ctypes.c_void_p()
        st = lib.tb3_create(int(device), ctypes.byref(ctx))
        if st != _abi.OK:
            msg = lib.tb3_last_error(None)
            raise TB3Error(st, msg.decode() if msg else "tb3_create failed")
        self._ctx = ctx
        self.n = 0
        self.algo = _abi.NATA
        self.nu = 1.0
        self.cutoff = 0.0
        self.precision = _abi.FP64_EXACT
        self.nrep = 1

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.tb3_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._ctx

    def _check(self, st):
        _abi.check(self._ctx, st)

    # -- configuration ------------------------------------------------------
    def configure(self, algo="nata", nu=1.0, cutoff=0.0, domain_L=None, precision="exact"):
        a = _abi.ALGORITHMS[algo] if isinstance(algo, str) else int(algo)
        p = _abi.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        L = None
        if domain_L is not None:
            L = _as(domain_L, (3,))
        self._check(self._lib.tb3_configure(self._ctx, a, float(nu), float(cutoff), _dp(L) if L is not None else None,
                                            p))
        self.algo, self.nu, self.cutoff, self.precision = a, float(nu), float(cutoff), p
        return self

    def set_particles(self, pos, vel=None, acc=None, mass=None, ids=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.zeros((n, 3)) if vel is None else _as(vel, (n, 3))
        acc = np.zeros((n, 3)) if acc is None else _as(acc, (n, 3))
        mass = np.ones(n) if mass is None else _as(mass, (n,))
        idp = None
        if ids is not None:
            ids = _as(ids, (n,), np.int32)
            idp = ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self._check(self._lib.tb3_set_particles(self._ctx, n, idp, _dp(pos), _dp(vel), _dp(acc), _dp(mass)))
        self.n = n
        self.nrep = 1
        return self

    def get_particles(self):
        n = self.n
        out = {"id": np.zeros(n, np.int32), "pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)),
               "acc": np.zeros((n, 3)), "mass": np.zeros(n), "f": np.zeros((n, 3))}
        self._check(self._lib.tb3_get_particles(
            self._ctx, n, out["id"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(out["pos"]),
            _dp(out["vel"]), _dp(out["acc"]), _dp(out["mass"]), _dp(out["f"])))
        return out

    # -- the hot path -------------------------------------------------------
    def step(self, dt, g=(0.0, 0.0, 0.0)):
        """One Simulation::Start iteration; returns (buffer triplets, particle triplets)."""
        gv = _as(g, (3,))
        nb, nt = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(self._lib.tb3_step(self._ctx, float(dt), _dp(gv), ctypes.byref(nb), ctypes.byref(nt)))
        return int(nb.value), int(nt.value)

    def run(self, steps, dt, g=(0.0, 0.0, 0.0)):
        gv = _as(g, (3,))
        counts = np.zeros(max(int(steps), 1), dtype=np.uint64)
        self._check(self._lib.tb3_run(self._ctx, int(steps), float(dt), _dp(gv),
                                      counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return counts[:steps]

    # -- replica batches -------------------------------------------------------
    def batch_set(self, offsets):
        """Split the particles into replicas: replica r holds rows offsets[r] .. offsets[r+1]-1."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self._check(self._lib.tb3_batch_set(self._ctx, len(off) - 1,
                                            off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.nrep = len(off) - 1
        return self

    def batch_run(self, steps, dt, g=(0.0, 0.0, 0.0)):
        """tb3_run with per-replica counts: an array of shape (steps, nrep)."""
        gv = _as(g, (3,))
        nrep = self.nrep
        counts = np.zeros((max(int(steps), 1), nrep), dtype=np.uint64)
        self._check(self._lib.tb3_batch_run(self._ctx, int(steps), float(dt), _dp(gv),
                                            counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return counts[:steps]

    def batch_counts(self):
        nrep = self.nrep
        counts = np.zeros(nrep, dtype=np.uint64)
        self._check(self._lib.tb3_batch_counts(self._ctx, nrep, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return counts

    def batch_energy(self):
        """Per-replica energies and triplet counts of the current positions."""
        nrep = self.nrep
        E = np.zeros(nrep)
        counts = np.zeros(nrep, dtype=np.uint64)
        self._check(self._lib.tb3_batch_energy(self._ctx, nrep, _dp(E),
                                               counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return E, counts

    def compute_forces(self):
        nt = ctypes.c_uint64(0)
        self._check(self._lib.tb3_compute_forces(self._ctx, ctypes.byref(nt)))
        return int(nt.value)

    def energy(self):
        E, nt = ctypes.c_double(0.0), ctypes.c_uint64(0)
        self._check(self._lib.tb3_energy(self._ctx, ctypes.byref(E), ctypes.byref(nt)))
        return E.value, int(nt.value)

    def energy_grad(self):
        E, dnu, nt = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_uint64(0)
        g = np.zeros((self.n, 3))
        self._check(self._lib.tb3_energy_grad(self._ctx, ctypes.byref(E), _dp(g), ctypes.byref(dnu),
                                              ctypes.byref(nt)))
        return E.value, g, dnu.value, int(nt.value)

    # -- device-pointer forms (torch) -----------------------------------------
    def energy_dev(self, n, d_pos, nu, d_E, stream=None, want_count=False):
        nt = ctypes.c_uint64(0)
        self._check(self._lib.tb3_energy_dev(self._ctx, int(n), ctypes.c_void_p(d_pos), float(nu),
                                             ctypes.c_void_p(d_E), ctypes.c_void_p(stream or 0),
                                             ctypes.byref(nt) if want_count else None))
        return int(nt.value) if want_count else None

    def energy_grad_dev(self, n, d_pos, nu, d_E, d_grad, stream=None, want_count=False):
        nt = ctypes.c_uint64(0)
        self._check(self._lib.tb3_energy_grad_dev(self._ctx, int(n), ctypes.c_void_p(d_pos), float(nu),
                                                  ctypes.c_void_p(d_E), ctypes.c_void_p(d_grad),
                                                  ctypes.c_void_p(stream or 0),
                                                  ctypes.byref(nt) if want_count else None))
        return int(nt.value) if want_count else None

    def energy_grad_dev_noalloc(self, n, d_pos, nu, d_E, d_grad, stream=None, want_count=False):
        """tb3_energy_grad_dev without device allocation (the library's
        static scratch array; P3BCA, fp64 or fp32-mixed)."""
        nt = ctypes.c_uint64(0)
        self._check(self._lib.tb3_energy_grad_dev_noalloc(self._ctx, int(n), ctypes.c_void_p(d_pos), float(nu),
                                                          ctypes.c_void_p(d_E), ctypes.c_void_p(d_grad),
                                                          ctypes.c_void_p(stream or 0),
                                                          ctypes.byref(nt) if want_count else None))
        return int(nt.value) if want_count else None

    def autotune(self, reps=3):
        """Pick the cell width / cutoff ratio, and the dense-owner kernel's
        waves per owner where that kernel runs, by timing force sweeps of the
        current particles; returns the best sweep time and every knob."""
        ms, ratio, blocks = ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_int(0)
        self._check(self._lib.tb3_autotune(self._ctx, int(reps), ctypes.byref(ms), ctypes.byref(ratio),
                                           ctypes.byref(blocks)))
        return {"sweep_ms": ms.value, **self.get_tuning()}

    def force_scale(self):
        """Each particle's force scale: the summed lengths of its triplet
        contributions (P3BCA with a cutoff), the yardstick of force errors."""
        out = np.zeros(self.n)
        self._check(self._lib.tb3_force_scale(self._ctx, _dp(out)))
        return out

    def get_tuning(self):
        """The knobs in force: cell width / cutoff ratio, the cell kernel's
        resident workgroups per CU, the dense-owner kernel's waves per owner."""
        ratio, blocks, waves = ctypes.c_double(0.0), ctypes.c_int(0), ctypes.c_int(0)
        self._check(self._lib.tb3_get_tuning(self._ctx, ctypes.byref(ratio), ctypes.byref(blocks),
                                             ctypes.byref(waves)))
        return {"cell_ratio": ratio.value, "blocks_per_cu": blocks.value, "dense_waves": waves.value}

    def set_tuning(self, cell_ratio=1.0, dense_waves=8):
        """Fix the knobs (cell_ratio in [1, 4]; dense_waves 4, 8 or 16)."""
        self._check(self._lib.tb3_set_tuning(self._ctx, float(cell_ratio), int(dense_waves)))

    def set_cell_slots(self, slots=-1):
        """Slots per cell of the slotted cell layout that replayed steps
        (run) use on the fast cutoff path: -1 automatic, 0 off, 1 .. 256."""
        self._check(self._lib.tb3_set_cell_slots(self._ctx, int(slots)))

    def get_cell_slots(self):
        """(requested, in_use): the setting, and the slots per cell of the
        last captured step graph (0: contiguous layout)."""
        req, use = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self._lib.tb3_get_cell_slots(self._ctx, ctypes.byref(req), ctypes.byref(use)))
        return req.value, use.value

    # -- phase timing ---------------------------------------------------------
    def set_profiling(self, enable=True):
        self._check(self._lib.tb3_set_profiling(self._ctx, 1 if enable else 0))

    def get_profile(self):
        """Accumulated phase times (HIP events on the engine's stream) since set_profiling."""
        v = np.zeros(6)
        self._check(self._lib.tb3_get_profile(self._ctx, _dp(v), 6))
        return {"steps": int(v[0]), "predict_ms": v[1], "neighbours_ms": v[2], "forces_ms": v[3],
                "update_ms": v[4], "triplets": int(v[5])}

    def selftest_fp64(self, a, b):
        """Device sqrt(a), a / b and the raw v_rsq_f64 estimate of a."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        sq, q, rs = np.zeros_like(a), np.zeros_like(a), np.zeros_like(a)
        self._check(self._lib.tb3_selftest_fp64(self._ctx, a.size, _dp(a), _dp(b), _dp(sq), _dp(q), _dp(rs)))
        return sq, q, rs

    def selftest_fp64_inrange(self, a, b):
        """The bit-exact path's in-range sequences: device sqrt(a) and a / b."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        sq, q = np.zeros_like(a), np.zeros_like(a)
        self._check(self._lib.tb3_selftest_fp64_inrange(self._ctx, a.size, _dp(a), _dp(b), _dp(sq), _dp(q)))
        return sq, q


# -- particle CSV (utility/utility.cpp:14-24, 54-68) ---------------------------

CSV_HEADER_IN = "ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m"
CSV_HEADER_OUT = "ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m, fX, fY, fZ"


def read_csv(path):
    """Particles from the reference CSV layout: (ids, pos, vel, acc, mass)."""
    ids, rows = [], []
    with open(path) as fh:
        fh.readline()
        for line in fh:
            line = line.strip()
            if not line:
                continue
            cells = line.split(",")
            if len(cells) < 11:
                raise ValueError(f"{path}: expected at least 11 columns")
            ids.append(int(cells[0]))
            rows.append([float(c) for c in cells[1:11]])
    r = np.array(rows, dtype=np.float64).reshape(-1, 10)
    return np.array(ids, dtype=np.int32), r[:, 0:3].copy(), r[:, 3:6].copy(), r[:, 6:9].copy(), r[:, 9].copy()


def write_csv(path, ids, pos, vel, acc, mass, f=None, full_precision=False):
    """Rows sorted by ID, ", " separators, %g (or %.17g) values."""
    fmt = "%.17g" if full_precision else "%g"
    order = np.argsort(np.asarray(ids), kind="stable")
    with open(path, "w") as fh:
        fh.write((CSV_HEADER_OUT if f is not None else CSV_HEADER_IN) + "\n")
        for i in order:
            vals = list(pos[i]) + list(vel[i]) + list(acc[i]) + [mass[i]]
            if f is not None:
                vals += list(f[i])
            fh.write(str(int(ids[i])) + "".join(", " + (fmt % v) for v in vals) + "\n")
