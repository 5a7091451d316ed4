#!/usr/bin/env python3
"""Regenerates the committed fixtures in tests/golden/.

  generator_draws.json  particle coordinates from libstdc++'s engines drawn in
                        the reference generator's pattern (gen_draws.cpp,
                        compiled here with g++), plus the standard's engine
                        known answers.
  oracle_golden.json    small seeded trajectories of the CPU oracle (inputs,
                        per-step counts and final state as hex floats): NATA,
                        AUTA and P3BCA; pins the oracle against regressions and
                        gives the GPU tests fixtures that need no oracle at
                        run time.

Run from anywhere: python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from threebodymdalgorithms_amd import generators  # noqa: E402
import oracle_lib  # noqa: E402

CASES = [
    # name, algo, n, box, cutoff, steps, seed
    ("nata_n10", "nata", 10, 2.5, 0.0, 1, 11),
    ("auta_n24", "auta", 24, 3.0, 0.0, 2, 7),
    ("p3bca_n80_c1.5", "p3bca", 80, 3.42, 1.5, 3, 5),
    ("p3bca_n300_c2", "p3bca", 300, 5.31, 2.0, 4, 13),
]


def hexs(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64).ravel()]


def generator_draws():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen_draws")
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "gen_draws.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    rec = {"compiler": cxx, "uniform": [], "clustered": [], "kat": None}
    for line in out.splitlines():
        f = line.split()
        if f[0] in ("uniform", "clustered"):
            rec[f[0]].append([float.fromhex(v).hex() for v in f[1:4]])
        elif f[0] == "kat":
            rec["kat"] = [int(f[1]), int(f[2])]
    return rec


def oracle_cases():
    out = []
    g = np.array([0.0, -9.81, 0.0])
    for name, algo, n, box, cutoff, steps, seed in CASES:
        P = generators.csv_rounded(generators.uniform(n, box, seed))
        V = np.zeros_like(P)
        A = np.zeros_like(P)
        M = np.ones(n)
        pos, vel, acc, f, counts = oracle_lib.simulate(P, V, A, M, algo, steps, 0.001, g, cutoff=cutoff,
                                                       mode="brute")
        out.append({"name": name, "algo": algo, "n": n, "cutoff": cutoff, "steps": steps, "dt": 0.001,
                    "g": list(g), "pos0": hexs(P), "counts": [int(c) for c in counts], "pos": hexs(pos),
                    "vel": hexs(vel), "acc": hexs(acc), "f": hexs(f)})
    rng = np.random.default_rng(99)
    trip = []
    for _ in range(6):
        T = rng.uniform(0.0, 2.0, size=(3, 3))
        ci, cj, ck = oracle_lib.at_contrib(T[0], T[1], T[2])
        trip.append({"pos": hexs(T), "ci": hexs(ci), "cj": hexs(cj), "ck": hexs(ck)})
    return {"trajectories": out, "triplets": trip}


def main():
    with open(os.path.join(HERE, "generator_draws.json"), "w") as fh:
        json.dump(generator_draws(), fh, indent=1)
    with open(os.path.join(HERE, "oracle_golden.json"), "w") as fh:
        json.dump(oracle_cases(), fh)
    print("wrote", os.path.join(HERE, "generator_draws.json"), os.path.join(HERE, "oracle_golden.json"))


if __name__ == "__main__":
    main()
