"""Measurements of the workload configurations beside bench.py's headline.

    python -m threebodymdalgorithms_amd.benchmarks auta2048 [--skip-exact]
    python -m threebodymdalgorithms_amd.benchmarks clustered16k [--skip-exact]

auta2048      AUTA, FP64, N=2048 uniform in a 10^3 box (generator seed
              926762934), mass 1, one step at dt 0.001 with g = (0,-9.81,0):
              exact, FP64 and FP32-mixed step times and unique triplets/s.
clustered16k  AUTA, one force sweep, N=16375 particles in 25 Gaussian
              clusters (centres uniform in 20^3 from seed 89347587, offsets
              N(0,1) from seed 926762934: the reference generator's
              numParticles / numClusters per cluster): FP32-mixed and FP64
              sweep times, and the FP32-mixed force and energy errors against
              the bit-exact FP64 sweep (which equals the reference's).

Each prints one JSON line.  Timings use HIP events recorded around the sweep
kernel on the engine's stream (tb3_set_profiling) and wall-clock time around
the synchronous step.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import generators
from .engine import Engine

FLOP_PER_TRIPLET = 115


def _time_step(eng, P, algo, precision, cutoff=0.0, reps=3):
    best = None
    for _ in range(reps):
        eng.set_particles(P)
        eng.configure(algo, cutoff=cutoff, precision=precision)
        eng.set_profiling(True)
        t0 = time.perf_counter()
        _, cnt = eng.step(0.001, (0.0, -9.81, 0.0))
        dt = time.perf_counter() - t0
        prof = eng.get_profile()
        eng.set_profiling(False)
        rec = {"ms_per_step": dt * 1e3, "force_kernel_ms": prof["forces_ms"], "triplets": cnt,
               "triplets_per_s": cnt / dt,
               "force_kernel_tflops": FLOP_PER_TRIPLET * cnt / (prof["forces_ms"] * 1e-3) / 1e12}
        if best is None or rec["ms_per_step"] < best["ms_per_step"]:
            best = rec
    return best


def _sweep(eng, P, precision):
    eng.set_particles(P)
    eng.configure("auta", precision=precision)
    eng.set_profiling(True)
    t0 = time.perf_counter()
    E, g, _, cnt = eng.energy_grad()
    dt = time.perf_counter() - t0
    prof = eng.get_profile()
    eng.set_profiling(False)
    return E, g, cnt, dt, prof


def auta2048(skip_exact=False):
    P = generators.csv_rounded(generators.uniform(2048, 10.0, 926762934))
    out = {"config": "AUTA N=2048 uniform 10^3, 1 step, dt 0.001, g (0,-9.81,0)"}
    with Engine(0) as eng:
        for prec in (["fp64", "fp32"] if skip_exact else ["exact", "fp64", "fp32"]):
            out[prec] = _time_step(eng, P, "auta", prec)
    return out


def clustered16k(skip_exact=False):
    P = generators.csv_rounded(generators.clustered_gauss(16384, 25, 20.0, 0.0, 1.0, 926762934, 89347587))
    out = {"config": f"AUTA one sweep, N={len(P)} in 25 Gaussian clusters (sd 1) in 20^3"}
    with Engine(0) as eng:
        E32, g32, c32, t32, _ = _sweep(eng, P, "fp32")
        E64, g64, c64, t64, _ = _sweep(eng, P, "fp64")
        out["fp32"] = {"sweep_s": t32, "triplets": c32, "triplets_per_s": c32 / t32}
        out["fp64"] = {"sweep_s": t64, "triplets": c64, "triplets_per_s": c64 / t64}
        out["fp32_speedup_vs_fp64"] = t64 / t32
        ref_g, ref_E, ref_name = g64, E64, "fp64"
        if not skip_exact:
            Ee, ge, ce, te, _ = _sweep(eng, P, "exact")
            out["exact"] = {"sweep_s": te, "triplets": ce}
            ref_g, ref_E, ref_name = ge, E64, "exact"
        num = np.linalg.norm(g32 - ref_g, axis=1)
        den = np.linalg.norm(ref_g, axis=1)
        rel = num / np.where(den > 0, den, 1.0)
        floor = np.sqrt(np.mean(den ** 2))
        out["fp32_vs_" + ref_name] = {"max_rel_force_err": float(rel.max()),
                                      "rms_rel_force_err": float(np.sqrt(np.mean(rel ** 2))),
                                      "max_rel_force_err_floored": float((num / np.maximum(den, floor)).max()),
                                      "rel_energy_err": float(abs(E32 - ref_E) / abs(ref_E))}
        num64 = np.linalg.norm(g64 - ref_g, axis=1)
        out["fp64_vs_" + ref_name] = {"max_rel_force_err": float((num64 / np.where(den > 0, den, 1.0)).max())}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["auta2048", "clustered16k"])
    ap.add_argument("--skip-exact", action="store_true")
    a = ap.parse_args(argv)
    fn = {"auta2048": auta2048, "clustered16k": clustered16k}[a.which]
    print(json.dumps(fn(skip_exact=a.skip_exact)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
