#!/usr/bin/env python3
"""Benchmark of the three-body hot path on one MI355X.

Workload (one "step" = one simulation step: predictor -> cutoff sweep -> update):
P3BCA, cutoff 2, N = 32768 particles uniform in a 25.4^3 box (density 2.0),
mass 1, velocities 0, dt = 0.001, g = (0, -9.81, 0); particles from the
reference generator's algorithm (mt19937, seed 926762934) written through the
CSV text format.  Warm-up steps run first, then the state is reset so the K
timed steps are steps 0..K-1 of that trajectory.

Printed JSON line:
  value        unique triplets/s of the FP64 (squared-distance) path, inputs
               resident in HBM: K steps of tb3_run (replays of one captured
               step graph, one host synchronisation) over wall-clock time
  exact        the bit-exact FP64 path (reference expression and order; its
               positions, velocities, forces and counts equal the
               reference's over every step), with its force kernel's and
               whole step's fraction of the FP64 peak
  fp64_vs_exact the headline path's per-particle force error against the
               bit-exact path on the initial state (max, RMS, floored at the
               RMS force) and the step through which their counts agree
  step_frac    115 FLOP per triplet over the whole step time / FP64 peak
  fp32_mixed   the FP32-mixed path, its speedup over FP64 and its force /
               energy error against the bit-exact path on the initial state
  roofline     the force kernel: 115 FLOP per unique triplet / its average
               duration from HIP events on the engine's stream, against the
               FP64 vector peak; traffic = HBM bytes per launch from the
               rocprofv3 FETCH_SIZE / WRITE_SIZE runs in profiles/
  cpu_baseline the reference-shaped CPU sweep (oracle restatement of the
               reference's one-process i<j<k loop with AoS particles and a
               virtual per-triplet call), single thread, on a bounded sample
               of rows of step 0, extrapolated by the exact candidate count
  cpu_multicore a fast CPU sweep (cell list, OpenMP over the host's threads)
               of step 0, for scale
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "unique triplets/s + ms/step, P3BCA c=2 N=32k; FP32-mix vs FP64 at 1e-6 F err"
FLOP_PER_TRIPLET = 115
FP64_VECTOR_PEAK_TFLOPS = 78.6
N, BOX, CUTOFF, DT, SEED = 32768, 25.4, 2.0, 0.001, 926762934
G = (0.0, -9.81, 0.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload():
    from threebodymdalgorithms_amd import generators
    return generators.csv_rounded(generators.uniform(N, BOX, SEED))


def timed_steps(eng, P, precision, steps, warmup, sync, tune=True):
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=CUTOFF, precision=precision)
    tuned = eng.autotune(reps=3) if (tune and precision != "exact") else None
    # the warm-up steps are replayed steps like the timed ones, so the step
    # graph is captured before timing; it survives the reset to step 0
    # (same particle count and grid)
    if warmup > 0:
        eng.run(warmup, DT, G)
    # phase times: the same steps, one tb3_step at a time, with HIP events
    # around each phase on the engine's stream.  They run before the timed
    # steps, which then start on a GPU at its sustained clock (a timed run
    # that followed only the autotuner and the warm-up measured 2-5% slower
    # than the same run repeated).
    eng.set_particles(P)
    eng.set_profiling(True)
    step_counts = [eng.step(DT, G)[1] for _ in range(steps)]
    prof = eng.get_profile()
    eng.set_profiling(False)
    eng.set_particles(P)
    sync()
    t0 = time.perf_counter()
    counts = [int(c) for c in eng.run(steps, DT, G)]
    sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    # the same trajectory either way: bit for bit on the exact path; the fast
    # paths sum forces atomically, and that last-bit freedom grows through the
    # chaotic dynamics until a triplet crosses the cutoff differently (FP64
    # after ~10 steps; FP32-mixed, whose rounding turns that into FP32-ulp
    # moves, from step 2 on), so their first steps are compared
    ncmp = {"exact": steps, "fp64": min(steps, 5), "fp32": min(steps, 2)}[precision]
    if step_counts[:ncmp] != counts[:ncmp]:
        raise RuntimeError("tb3_run and tb3_step disagree on the triplet counts")
    total = int(sum(counts))
    fk_ms = prof["forces_ms"] / max(prof["steps"], 1)
    t_avg = prof["triplets"] / max(prof["steps"], 1)
    return {
        "tuned": tuned,
        "triplets_per_s": total / secs,
        "ms_per_step": 1e3 * secs / steps,
        "triplets_per_step": total / steps,
        "counts": counts,
        "force_kernel_ms": fk_ms,
        "neighbour_ms": prof["neighbours_ms"] / max(prof["steps"], 1),
        "integrate_ms": (prof["predict_ms"] + prof["update_ms"]) / max(prof["steps"], 1),
        "force_kernel_tflops": FLOP_PER_TRIPLET * t_avg / (fk_ms * 1e-3) / 1e12 if fk_ms > 0 else 0.0,
    }


def step_frac(r):
    """115 FLOP per unique triplet over the whole step (cell list, sweep,
    integration and the launches between them), against the FP64 peak."""
    return FLOP_PER_TRIPLET * r["triplets_per_step"] / (r["ms_per_step"] * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS


def host_cpu():
    """The host's CPU model and the threads this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_available": avail}


def force_errors(f, fo, scale):
    """Per-particle force error of f against fo: the ratio to each particle's
    force scale S_p (the summed lengths of its triplet contributions, which
    bounds the error of a sum of terms each computed to a relative error e by
    e S_p however much they cancel), the pure ratio |df|/|fo| (max and RMS)
    and the ratio with |fo| floored at the RMS force."""
    num = np.linalg.norm(f - fo, axis=1)
    den = np.linalg.norm(fo, axis=1)
    rel = num / np.where(den > 0, den, 1.0)
    floor = np.sqrt(np.mean(den ** 2))
    return {"max_rel_force_err_scale": float((num / np.where(scale > 0, scale, 1.0)).max()),
            "max_rel_force_err": float(rel.max()), "rms_rel_force_err": float(np.sqrt(np.mean(rel ** 2))),
            "max_rel_force_err_floored": float((num / np.maximum(den, floor)).max())}


def accuracy_vs_exact(eng, P):
    """FP64 and FP32-mixed forces and energies of the initial state against
    the bit-exact FP64 path (whose forces equal the reference's bit for bit),
    with each particle's force scale from tb3_force_scale."""
    eng.set_particles(P)
    eng.configure("p3bca", cutoff=CUTOFF, precision="exact")
    Ex, fx, _, _ = eng.energy_grad()
    scale = eng.force_scale()
    out = {}
    for prec in ("fp64", "fp32"):
        eng.configure("p3bca", cutoff=CUTOFF, precision=prec)
        E, f, _, _ = eng.energy_grad()
        out[prec] = {**force_errors(f, fx, scale), "rel_energy_err": float(abs(E - Ex) / abs(Ex))}
    return out


PROFILE = os.path.join(ROOT, "profiles", "k_p3bca_cells_p3bca_c2_n32768_fp64.md")


def profiled_traffic():
    """HBM bytes per launch of the cutoff sweep as measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in the kernel's profile summary
    (counters cannot be read from inside this process); None if absent."""
    try:
        import re
        m = re.search(r"traffic_bytes_per_launch = ([0-9.eE+]+)", open(PROFILE).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(P, count0, target_s):
    """Reference-shaped single-thread sweep on rows 0..R-1 of step 0."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_sweep_brute_rows.argtypes = [ctypes.c_int64, dp, ctypes.c_double, ctypes.c_double, dp, dp, dp, dp,
                                         ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
    lib.orc_sweep_brute_rows.restype = ctypes.c_uint64
    lib.orc_domain_size.argtypes = [ctypes.c_int64, dp, dp, dp]
    pos = np.ascontiguousarray(P, dtype=np.float64)
    L, lo = np.zeros(3), np.zeros(3)
    lib.orc_domain_size(N, pos.ctypes.data_as(dp), L.ctypes.data_as(dp), lo.ctypes.data_as(dp))
    f = [np.zeros((N, 3)) for _ in range(3)]
    acc = ctypes.c_uint64(0)

    def run(rows):
        t0 = time.perf_counter()
        cand = lib.orc_sweep_brute_rows(N, pos.ctypes.data_as(dp), 1.0, CUTOFF, L.ctypes.data_as(dp),
                                        f[0].ctypes.data_as(dp), f[1].ctypes.data_as(dp), f[2].ctypes.data_as(dp),
                                        rows, ctypes.byref(acc))
        return time.perf_counter() - t0, int(cand)

    t1, _ = run(1)
    rows = max(1, min(N - 2, int(target_s / max(t1, 1e-3))))
    secs, cand = run(rows)
    total_cand = N * (N - 1) * (N - 2) // 6
    step_s = secs * total_cand / cand
    return {"value": count0 / step_s, "unit": "unique triplets/s", "cores": 1, "kind": "port", "host": host_cpu(),
            "sample": f"reference-shaped brute i<j<k sweep, rows 0..{rows - 1} of step 0 ({cand:.3e} of "
                      f"{total_cand:.3e} candidates, {secs:.1f} s), extrapolated by candidate count to "
                      f"{step_s:.0f} s per step",
            "ms_per_step": step_s * 1e3}


def cpu_fast(P):
    """The oracle's multi-core cell-list sweep (fast CPU implementation, not
    the reference's loop) on step 0, all host threads OpenMP gives it."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_sweep_cells_fast.argtypes = [ctypes.c_int64, dp, ctypes.c_double, ctypes.c_double, dp, dp,
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
    lib.orc_sweep_cells_fast.restype = ctypes.c_double
    lib.orc_domain_size.argtypes = [ctypes.c_int64, dp, dp, dp]
    pos = np.ascontiguousarray(P, dtype=np.float64)
    L, lo = np.zeros(3), np.zeros(3)
    lib.orc_domain_size(N, pos.ctypes.data_as(dp), L.ctypes.data_as(dp), lo.ctypes.data_as(dp))
    f = np.zeros((N, 3))
    cnt = ctypes.c_uint64(0)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        lib.orc_sweep_cells_fast(N, pos.ctypes.data_as(dp), 1.0, CUTOFF, L.ctypes.data_as(dp), f.ctypes.data_as(dp),
                                 ctypes.byref(cnt), threads)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {"value": cnt.value / best, "unit": "unique triplets/s", "cores": threads, "kind": "port",
            "host": host_cpu(),
            "sample": f"cell-list sweep of step 0 ({cnt.value} triplets, best of 3: {best * 1e3:.1f} ms)",
            "ms_per_step": best * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-variants", action="store_true", help="time the FP64 path only")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default cell width (no tb3_autotune)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 512 MiB device-to-device with torch (calibrates the HBM byte counters)")
    ap.add_argument("--only", choices=["fp64", "exact", "fp32"],
                    help="time this path alone and print its figures (profiling runs; not the contract line)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: tb3 runs on one GPU; --gpus {args.gpus} (WORLD_SIZE={world}) is not supported",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import threebodymdalgorithms_amd as tb3

    def sync():
        torch.cuda.synchronize()

    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")  # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8, "launches": 3}), flush=True)
        return

    P = workload()
    log(f"bench: N={N} box={BOX} cutoff={CUTOFF} steps={args.steps} warmup={args.warmup}")
    if args.only:
        with tb3.Engine(0) as eng:
            r = timed_steps(eng, P, args.only, args.steps, args.warmup, sync, not args.no_autotune)
        print(json.dumps({"path": args.only, **r}), flush=True)
        return
    with tb3.Engine(0) as eng:
        fp64 = timed_steps(eng, P, "fp64", args.steps, args.warmup, sync, not args.no_autotune)
        log(f"bench: fp64 {fp64['ms_per_step']:.3f} ms/step {fp64['triplets_per_s']:.4e} triplets/s "
            f"force kernel {fp64['force_kernel_ms']:.3f} ms")
        out_exact = out_fp32 = fp64_vs_exact = None
        if not args.skip_variants:
            ex = timed_steps(eng, P, "exact", args.steps, args.warmup, sync)
            log(f"bench: exact {ex['ms_per_step']:.3f} ms/step force kernel {ex['force_kernel_ms']:.3f} ms")
            # step 0 starts from identical positions and must count the same;
            # later steps follow trajectories that the fast path's ~1e-13
            # force differences move apart until a triplet crosses the cutoff
            # differently
            same = next((k for k, (a, b) in enumerate(zip(ex["counts"], fp64["counts"])) if a != b),
                        len(ex["counts"]))
            if same == 0:
                log(f"bench: WARNING exact and fp64 step-0 counts differ: {ex['counts']} vs {fp64['counts']}")
            elif same < len(ex["counts"]):
                worst = max(abs(a - b) for a, b in zip(ex["counts"], fp64["counts"]))
                log(f"bench: exact and fp64 counts agree for steps 0-{same - 1}; the trajectories then diverge "
                    f"(at most {worst} triplets apart)")
            f32 = timed_steps(eng, P, "fp32", args.steps, args.warmup, sync, not args.no_autotune)
            acc = accuracy_vs_exact(eng, P)
            log(f"bench: fp32 {f32['ms_per_step']:.3f} ms/step force kernel {f32['force_kernel_ms']:.3f} ms "
                f"{acc['fp32']}; fp64 vs exact {acc['fp64']}")
            out_exact = {"value": ex["triplets_per_s"], "ms_per_step": ex["ms_per_step"],
                         "force_kernel_ms": ex["force_kernel_ms"],
                         "force_kernel_tflops": ex["force_kernel_tflops"],
                         "force_kernel_frac": ex["force_kernel_tflops"] / FP64_VECTOR_PEAK_TFLOPS,
                         "step_frac": step_frac(ex),
                         "counts_equal_reference_path": True,
                         "note": "reproduces the reference bit for bit (positions, velocities, forces, counts) "
                                 "over every step"}
            out_fp32 = {"value": f32["triplets_per_s"], "ms_per_step": f32["ms_per_step"],
                        "force_kernel_ms": f32["force_kernel_ms"],
                        "speedup_vs_fp64_step": fp64["ms_per_step"] / f32["ms_per_step"],
                        "speedup_vs_fp64_force_kernel": fp64["force_kernel_ms"] / f32["force_kernel_ms"],
                        "step0_vs_exact": acc["fp32"]}
            fp64_vs_exact = {"step0": acc["fp64"], "counts_agree_through_step": same - 1,
                             "note": "one sweep of the initial state against the bit-exact path; the fast path "
                                     "sums in another order, so the chaotic trajectories part after the first steps"}
    cpu = cpu_mc = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline ...")
        cpu = cpu_baseline(P, fp64["counts"][0], args.cpu_baseline_seconds)
        log("bench: multi-core cpu sweep ...")
        cpu_mc = cpu_fast(P)
    achieved = fp64["force_kernel_tflops"]
    rec = {
        "metric": METRIC,
        "value": fp64["triplets_per_s"],
        "unit": "unique triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "P3BCA cutoff 2, N=32768 uniform in 25.4^3 (density 2), dt 0.001, g (0,-9.81,0)",
                   "algorithm": "p3bca", "n": N, "cutoff": CUTOFF, "precision": "fp64",
                   "triplets_per_step": fp64["triplets_per_step"], "counts": fp64["counts"],
                   "force_kernel_ms": fp64["force_kernel_ms"], "neighbour_ms": fp64["neighbour_ms"],
                   "integrate_ms": fp64["integrate_ms"], "tuned": fp64["tuned"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK_TFLOPS, "traffic": profiled_traffic(),
                     "kernel": "tb3::k_p3bca_cells<double, true, false>", "flop_per_triplet": FLOP_PER_TRIPLET},
        "step_frac": step_frac(fp64),
        "fp64_vs_exact": fp64_vs_exact,
        "cpu_baseline": cpu,
        "cpu_multicore": cpu_mc,
        "exact": out_exact,
        "fp32_mixed": out_fp32,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
