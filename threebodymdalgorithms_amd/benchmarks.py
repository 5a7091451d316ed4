"""Measurements of the workload configurations beside bench.py's headline.

    python -m threebodymdalgorithms_amd.benchmarks auta2048 [--skip-exact]
    python -m threebodymdalgorithms_amd.benchmarks clustered16k [--skip-exact]
    python -m threebodymdalgorithms_amd.benchmarks relax8k
    python -m threebodymdalgorithms_amd.benchmarks batch256 [--steps K] [--skip-exact]

auta2048      AUTA, FP64, N=2048 uniform in a 10^3 box (generator seed
              926762934), mass 1, one step at dt 0.001 with g = (0,-9.81,0):
              exact, FP64 and FP32-mixed step times and unique triplets/s.
clustered16k  AUTA, one force sweep, N=16375 particles in 25 Gaussian
              clusters (centres uniform in 20^3 from seed 89347587, offsets
              N(0,1) from seed 926762934: the reference generator's
              numParticles / numClusters per cluster): FP32-mixed and FP64
              sweep times, and the FP32-mixed force and energy errors against
              the bit-exact FP64 sweep (which equals the reference's).
relax8k       P3BCA cutoff 2.5, FP64, N=8000: a 20^3 grid of spacing 0.8
              (16^3 box) jittered by N(0, 0.05) per coordinate (numpy PCG64,
              seed 0 + configuration number).  Times one energy forward +
              adjoint through torch.autograd, runs 200 steps of gradient
              descent (step 1e-5), and fits nu by force matching (L-BFGS)
              over 32 configurations whose target forces come from the
              bit-exact path at nu_true = 1.3.
batch256      P3BCA cutoff 2, FP32-mixed: 256 replicas of N log-uniform in
              [256, 4096] (numpy PCG64 seed 926762934), each one Gaussian
              cloud N(0, 1.5) per axis (the reference's Gauss generator,
              minstd_rand0 seed 1000 + r), written through the CSV text
              format, each on its own bbox-rule periodic domain; K steps (100)
              of tb3_batch_run, beside FP64, and the FP32-mixed force and
              energy errors against the bit-exact path on the initial state.

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


def jittered_grid(per_dim=20, spacing=0.8, sigma=0.05, seed=0):
    g = np.arange(per_dim) * spacing
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    return P + np.random.default_rng(seed).normal(0.0, sigma, P.shape)


def relax8k(skip_exact=False):
    import torch
    from .autograd import at_energy, at_forces
    cutoff, L = 2.5, np.full(3, 16.0)
    out = {"config": "P3BCA c=2.5 FP64, N=8000 jittered 20^3 grid (spacing 0.8, sd 0.05) in 16^3"}
    with Engine(0) as eng:
        P = jittered_grid(seed=0)
        eng.set_particles(P)
        eng.configure("p3bca", nu=1.0, cutoff=cutoff, domain_L=L, precision="fp64")
        x = torch.tensor(P, dtype=torch.float64, device="cuda", requires_grad=True)
        nu = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
        for _ in range(2):
            x.grad = None
            at_energy(x, nu, eng).backward()
        torch.cuda.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            x.grad = None
            nu.grad = None
            at_energy(x, nu, eng).backward()
        torch.cuda.synchronize()
        out["forward_backward_ms"] = (time.perf_counter() - t0) / reps * 1e3
        _, cnt = eng.energy()
        out["triplets"] = cnt

        energies = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(200):
            x.grad = None
            E = at_energy(x, 1.0, eng)
            E.backward()
            with torch.no_grad():
                x.copy_(x - 1e-5 * x.grad)
            energies.append(E.detach())
        torch.cuda.synchronize()
        e = torch.stack(energies).cpu().numpy()
        out["relax"] = {"steps": 200, "step_size": 1e-5, "s": time.perf_counter() - t0, "E0": float(e[0]),
                        "E_final": float(e[-1]), "monotone": bool(np.all(np.diff(e) < 0))}

        nu_true = 1.3
        configs, targets = [], []
        t0 = time.perf_counter()
        for s in range(32):
            Q = jittered_grid(seed=s)
            eng.set_particles(Q)
            eng.configure("p3bca", nu=nu_true, cutoff=cutoff, domain_L=L, precision="exact")
            _, f, _, _ = eng.energy_grad()
            configs.append(torch.tensor(Q, dtype=torch.float64, device="cuda"))
            targets.append(torch.tensor(f, dtype=torch.float64, device="cuda"))
        t_targets = time.perf_counter() - t0
        eng.configure("p3bca", nu=1.0, cutoff=cutoff, domain_L=L, precision="fp64")
        nu = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
        opt = torch.optim.LBFGS([nu], lr=1.0, max_iter=50, tolerance_grad=1e-14, tolerance_change=1e-16,
                                line_search_fn="strong_wolfe")
        evals = [0]

        def closure():
            opt.zero_grad()
            loss = sum(((at_forces(c, nu, eng) - t) ** 2).sum() for c, t in zip(configs, targets))
            loss.backward()
            evals[0] += 1
            return loss

        t0 = time.perf_counter()
        opt.step(closure)
        torch.cuda.synchronize()
        out["nu_fit"] = {"configs": 32, "nu_true": nu_true, "nu_fit": float(nu),
                         "rel_err": abs(float(nu) - nu_true) / nu_true, "loss_evals": evals[0],
                         "fit_s": time.perf_counter() - t0, "targets_s": t_targets}
    return out


def batch256(skip_exact=False, steps=100):
    rng = np.random.default_rng(926762934)
    sizes = np.rint(np.exp(rng.uniform(np.log(256), np.log(4096), 256))).astype(int)
    reps = [generators.csv_rounded(generators.gauss(int(n), 0.0, 1.5, 1000 + r)) for r, n in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    P = np.concatenate(reps)
    g = (0.0, -9.81, 0.0)
    out = {"config": f"P3BCA c=2, 256 Gaussian replicas (sd 1.5), N log-uniform [256, 4096], total {len(P)}",
           "n_total": int(len(P)), "n_min": int(sizes.min()), "n_max": int(sizes.max())}
    with Engine(0) as eng:
        for prec in ("fp32", "fp64"):
            eng.set_particles(P)
            eng.configure("p3bca", cutoff=2.0, precision=prec)
            eng.batch_set(off)
            # the cell width and the dense-owner kernel's waves per owner
            tuned = eng.autotune(reps=2)
            eng.batch_run(2, 0.001, g)
            # the force sweep of the initial state (the densest), then K steps
            eng.set_particles(P)
            eng.batch_set(off)
            eng.set_profiling(True)
            c0 = eng.compute_forces()
            sweep_ms = eng.get_profile()["forces_ms"]
            eng.set_profiling(False)
            eng.set_particles(P)
            eng.batch_set(off)
            t0 = time.perf_counter()
            counts = eng.batch_run(steps, 0.001, g)
            secs = time.perf_counter() - t0
            total = int(counts.sum())
            out[prec] = {"steps": steps, "ms_per_step": secs / steps * 1e3, "triplets_per_step": total / steps,
                         "triplets_per_s": total / secs, "step0_sweep_ms": sweep_ms, "step0_triplets": c0,
                         "step0_sweep_triplets_per_s": c0 / (sweep_ms * 1e-3), "tuned": tuned}
        out["fp32_speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out["fp32"]["ms_per_step"]
        out["fp32_sweep_speedup_vs_fp64"] = out["fp64"]["step0_sweep_ms"] / out["fp32"]["step0_sweep_ms"]
        if not skip_exact:
            eng.set_particles(P)
            eng.configure("p3bca", cutoff=2.0, precision="exact")
            eng.batch_set(off)
            Ee, ge, _, ce = eng.energy_grad()
            Eex, _ = eng.batch_energy()
            eng.configure("p3bca", cutoff=2.0, precision="fp32")
            eng.batch_set(off)
            E32, g32, _, c32 = eng.energy_grad()
            E32r, _ = eng.batch_energy()
            num = np.linalg.norm(g32 - ge, axis=1)
            den = np.linalg.norm(ge, axis=1)
            rel = num / np.where(den > 0, den, 1.0)
            floor = np.sqrt(np.mean(den ** 2))
            out["fp32_vs_exact"] = {"max_rel_force_err": float(rel.max()),
                                    "rms_rel_force_err": float(np.sqrt(np.mean(rel ** 2))),
                                    "max_rel_force_err_floored": float((num / np.maximum(den, floor)).max()),
                                    "max_rel_energy_err_per_replica":
                                        float(np.max(np.abs(E32r - Eex) / np.maximum(np.abs(Eex), 1e-300))),
                                    "counts_equal": bool(ce == c32)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["auta2048", "clustered16k", "relax8k", "batch256"])
    ap.add_argument("--skip-exact", action="store_true")
    ap.add_argument("--steps", type=int, default=100)
    a = ap.parse_args(argv)
    if a.which == "batch256":
        res = batch256(skip_exact=a.skip_exact, steps=a.steps)
    else:
        fn = {"auta2048": auta2048, "clustered16k": clustered16k, "relax8k": relax8k}[a.which]
        res = fn(skip_exact=a.skip_exact)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
