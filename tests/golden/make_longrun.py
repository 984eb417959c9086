#!/usr/bin/env python3
"""Writes tests/golden/longrun_n16384.npz: long-run flock statistics of the
reference's own dynamics, the pin of tests/test_gpu_longrun.py.

Workload: N = 16384 boids, F = 1 (the reference's boidsim_grid at density
16384, configuration 1's shape), the reference's random_uniform init at
srand48(seed) (src/boid_simulation_grid.cpp:95-136), T = 2000 updates, two
rule sets:
  cfg      radii 0.02/0.05/0.08, view 270 deg, weights 0.002/0.06/0.008,
           |v| in [0.001, 0.005] (disordered: polarisation ~ 1/sqrt(N));
  ordered  alignment radius 0.08 without view cone and weight 0.3,
           |v| in [0.004, 0.005] (the flock orders: polarisation ~ 0.9).
Every `every` updates the metrics of the state (bo_metrics: polarisation,
mean nearest-neighbour distance with the periodic metric of
tools/calc_distance.pyx:6-14, clusters of the graph d < 0.02 with >= 2
members, largest cluster) are recorded.

The dynamics are the float oracle's (oracle/boids_oracle_step.inc, float
instantiation, cell enumeration), which is bit-identical to the reference's
all-pairs and grid engines (golden fixtures of make_golden.py).  With
--check-ref the first seed of the cfg rule set is also run through the
reference's own grid engine (oracle/_ref/ref_grid, built from
src/boid_simulation_grid.cpp by oracle/Makefile.ref) and its end state must
equal the oracle's bit for bit, so the time series are the reference's.

TEST-FIXTURE GENERATOR: run in the build container (the reference tree and
its compiled harness are not on the GPU box); commit the .npz it writes.
"""
from __future__ import annotations

import argparse
import hashlib
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

RULESETS = {
    "cfg": dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001),
    "ordered": dict(sep=(0.02, 0.75, 0.002), ali=(0.08, 1.0, 0.3), coh=(0.08, 0.75, 0.008), vmax=0.005,
                    vmin=0.004),
}
EPS, MIN_SIZE = 0.02, 2


def run(rules, n, F, seed, T, every, threads):
    P = O.params(F, **rules)
    p, v = O.init_reference("random_uniform", n, F, rules["vmax"], rules["vmin"], seed=seed)
    series = []
    for t in range(T + 1):
        if t % every == 0:
            m = O.metrics(F, p.astype(np.float64), v.astype(np.float64), EPS, MIN_SIZE, threads)
            series.append([t, m["polarisation"], m["mean_nn"], m["clusters"], m["largest"]])
        if t < T:
            O.step(P, p, v, enumeration=O.ENUM_CELLS, nthreads=threads)
    return np.array(series, np.float64), p, v


def ref_grid_end_state(rules, n, F, seed, T):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_grid")
    fd, out = tempfile.mkstemp(suffix=".dump")
    os.close(fd)
    try:
        args = [exe, out, str(n), str(T), repr(F), *(repr(x) for x in (*rules["sep"], *rules["ali"], *rules["coh"])),
                repr(rules["vmax"]), repr(rules["vmin"]), "random_uniform", str(seed), str(T)]
        subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
        return O.read_ref_dump(out, n)[T]
    finally:
        os.unlink(out)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--field", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--seeds", default="1,2,3,4,5,6")
    ap.add_argument("--threads", type=int, default=len(os.sched_getaffinity(0)))
    ap.add_argument("--check-ref", action="store_true")
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "longrun_n16384.npz"))
    a = ap.parse_args()
    seeds = [int(s) for s in a.seeds.split(",")]
    out = {"n": np.int64(a.n), "field": np.float64(a.field), "steps": np.int64(a.steps),
           "every": np.int64(a.every), "seeds": np.array(seeds, np.int64), "eps": np.float64(EPS),
           "min_size": np.int64(MIN_SIZE)}
    for name, rules in RULESETS.items():
        out[f"{name}_rules"] = np.array([*rules["sep"], *rules["ali"], *rules["coh"], rules["vmax"], rules["vmin"]])
        allser, ends = [], []
        for seed in seeds:
            t0 = time.time()
            ser, p, v = run(rules, a.n, a.field, seed, a.steps, a.every, a.threads)
            allser.append(ser)
            ends.append(hashlib.sha256(p.tobytes() + v.tobytes()).hexdigest())
            print(f"{name} seed {seed}: {time.time() - t0:.0f} s; t={a.steps}: polarisation {ser[-1, 1]:.4f}, "
                  f"mean NN {ser[-1, 2]:.5f}, clusters {ser[-1, 3]:.0f}", flush=True)
            if a.check_ref and name == "cfg" and seed == seeds[0]:
                t0 = time.time()
                rp, rv, _ = ref_grid_end_state(rules, a.n, a.field, seed, a.steps)
                same = np.array_equal(rp, p) and np.array_equal(rv, v)
                print(f"  reference grid engine, {a.steps} updates ({time.time() - t0:.0f} s): end state "
                      f"{'bit-identical' if same else 'DIFFERS'}", flush=True)
                if not same:
                    sys.exit(1)
                out["ref_checked"] = np.array([name, str(seed), str(a.steps)])
        out[f"{name}_series"] = np.stack(allser)           # seeds x samples x (t, pol, nn, clusters, largest)
        out[f"{name}_end_sha256"] = np.array(ends)
    np.savez_compressed(a.out, **out)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
