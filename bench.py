#!/usr/bin/env python3
"""Boid-step throughput on one MI355X: FP64 plain vs mixed-precision grid path.

Workload (default, flock1m): N = 2^20 boids, cubic field F = 4.0 (density
16384 per unit volume), uniform positions, isotropic directions with |v| ~
U[0.001, 0.005] (seed 2), rules: radii 0.02/0.05/0.08 (separation/alignment/
cohesion), view angle 270 deg, weights 0.002/0.06/0.008, dt = 1, periodic
wrap, T = 1000 updates.  A "step" is one full update() of all N boids: fused
neighbour + integrate kernel, then the counting sort that re-bins the new
positions.  Cohesion clusters the flock as the run goes on (mean cohesion
neighbours 35 at t = 0, ~118 at t = 1000), so a step's cost depends on where
in the run it is.

Timed steps.  With --steps K >= T the timed region is the whole run from its
initial state (K consecutive updates).  With K < T the run is stratified:
the engine steps through the configured T updates once, and the K timed
updates are the ones at t_k = floor((k + 1/2) T / K), k = 0..K-1 (the
midpoints of K equal segments of the run); every timed update is bracketed
by device synchronisations, the updates between them run untimed.  `value`
= N K / (sum of the K timed update times): an estimate of the T-update run's
throughput.  The first K updates from t = 0 are timed as well and reported
as `from_t0` (the uniform start, the cheapest part of the run).

Prints ONE JSON line.  `value` is the headline path's boid-steps/s with the
state resident in HBM (flock1m: mixed precision; the FP64 plain figure
beside it).  `roofline` is for the dominant kernel (fused neighbour +
integrate) of the headline path, from HIP events on the engine's stream
over the timed updates (`traffic` and `valu` are read from the committed
rocprofv3 summaries); `cpu_baseline` times the CPU path beside it on this
host for a bounded sample, with all host threads and with one.

Other workloads (--workload): flock16k (N=16384, F=1, T=8000; headline:
the float grid mode that reproduces the reference's bytes, also FP64 plain
and mixed; the CPU beside it is the reference's own grid engine built from
source,
oracle/_ref/ref_grid), flock1k (N=1024, F=1, T=100, all-pairs FP64 and the
reference-faithful float mode; CPU: the reference's own all-pairs engine,
oracle/_ref/ref_allpairs) and blobs16m (N=2^24, F=10, 64 clustered blobs,
T=500, mixed).

`sweep` (with --sweep-steps > 0) reports the many-flock parameter sweep: 512
flocks, N log-uniform in [2^10, 2^16], field (N/16384)^(1/3), vmax ~
U[0.002, 0.01], vmin = vmax/5, radii x U[0.7, 1.3], mixed precision, seed 3,
run as one batched engine with graph-replayed steps, next to the same
kernels launched flock by flock (one engine per flock: flocks one after
another as separate jobs, eager or graph-replayed, and all flocks
interleaved on their own streams).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import boidssimulatormpi_amd as B  # noqa: E402
from boidssimulatormpi_amd import engine as E  # noqa: E402

RULES = dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001)
# Workloads (all with RULES): the headline (default) and three more shapes of
# the same path; `steps` is the run's configured length T.
WORKLOADS = {
    "flock1m": dict(n=1 << 20, field=4.0, seed=2, init=E.INIT_ISOTROPIC, steps=1000,
                    modes="grid_f64,grid_mixed", head="grid_mixed", cpu="port", sweep=True,
                    what="uniform positions, isotropic velocities"),
    "flock16k": dict(n=16384, field=1.0, seed=1, init=E.INIT_ISOTROPIC, steps=8000,
                     modes="grid_f32ref,grid_f64,grid_mixed", head="grid_f32ref", cpu="ref_grid", sweep=False,
                     what="uniform positions, isotropic velocities"),
    "flock1k": dict(n=1024, field=1.0, seed=0, init=E.INIT_ISOTROPIC, steps=100,
                    modes="allpairs_f64,allpairs_f32ref", head="allpairs_f64", cpu="ref_allpairs", sweep=False,
                    what="uniform positions, isotropic velocities"),
    "blobs16m": dict(n=1 << 24, field=10.0, seed=4, init=E.INIT_CLUSTERED, steps=500,
                     modes="grid_mixed", head="grid_mixed", cpu=None, sweep=False,
                     what="64 Gaussian blobs (sigma 0.05 F), per-blob headings"),
}
BYTES_PER_BOID_STEP = 96          # read + write of the FP64 master state (pos + vel, 6 x 8 B)
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def mid_run_state(init, n, field, t):
    """The workload's state after t updates (FP64 grid engine): a clustered
    state typical of the run, for tuning and for the CPU sample."""
    e = B.Engine(B.make_params(n, field, **RULES, mode=E.MODE_GRID_F64))
    e.set_state(xyzv6=init)
    e.step(t)
    pos, vel = e.state_f64()
    e.close()
    return pos, vel


def tuned_params(mode, mid, n, field, tune, block=0):
    """Engine parameters of one mode: the autotuner's choice, timed on the
    mid-run state `mid` = (pos, vel), or the defaults."""
    p = B.make_params(n, field, **RULES, mode=mode, flags=E.FLAG_TIMING, block=block)
    if not tune:
        return p, None
    from boidssimulatormpi_amd import autotune as T
    best, ps, timings = T.tune(p, {"pos": mid[0], "vel": mid[1]}, steps=6, warmup=2, cache=False,
                               blocks=(block,) if block else T.BLOCKS, fills=(100,))
    return ps[0], {"best": best, "candidates": len(timings),
                   "best_ms": min(t["ms_per_step"] for t in timings),
                   "worst_ms": max(t["ms_per_step"] for t in timings)}


def sample_times(T, K):
    """Timed updates of a stratified run: the midpoints of K equal segments
    of the T-update run (all of it, in order, when K >= T)."""
    if K >= T:
        return list(range(K))
    return [int((k + 0.5) * T / K) for k in range(K)]


def run_mode(mode, init, n, field, T, steps, warmup, params):
    """Times `steps` updates of one mode over the T-update run from `init`
    (sample_times), then the first `steps` updates from t = 0."""
    e = B.Engine(params)
    info = e.info()
    def load():
        if isinstance(init, tuple):
            e.set_state(pos=init[0], vel=init[1])
        else:
            e.set_state(xyzv6=init)
    load()
    e.step(warmup)
    load()                           # the run starts from the initial state
    e.sync()
    e.reset_timing()
    ts = sample_times(T, steps)
    pair0 = (0.0, 0)
    wall, pair_ms, pair_n, samples = 0.0, 0.0, 0, []
    t_cur = 0
    if steps >= T:                   # the whole run, consecutive updates
        t0 = time.perf_counter()
        e.step(steps)
        e.sync()
        wall = time.perf_counter() - t0
        pair_ms, pair_n = e.timing(E.K_PAIR)
        t_cur = steps
    else:
        for t in ts:
            if t > t_cur:
                e.step(t - t_cur)    # untimed
            e.sync()
            pair0 = e.timing(E.K_PAIR)
            t0 = time.perf_counter()
            e.step(1)
            e.sync()
            dt = time.perf_counter() - t0
            p1 = e.timing(E.K_PAIR)
            wall += dt
            pair_ms += p1[0] - pair0[0]
            pair_n += p1[1] - pair0[1]
            samples.append([t, round(1e3 * dt, 4), round(p1[0] - pair0[0], 4)])
            t_cur = t + 1
    kernels = {}
    for k, name in enumerate(E.K_NAMES):
        ms, cnt = e.timing(k)
        if cnt:
            kernels[name] = {"ms_total": ms, "launches": cnt, "us_avg": 1e3 * ms / cnt}
    # physical sanity of the state reached (cheap, outside the timed region)
    p, v = e.state_f64()
    sp = np.linalg.norm(v, axis=1)
    rel = 1e-6 if "f32" in E.MODE_NAMES[mode] else 1e-9     # float modes clamp |v| in float
    ok = bool(np.isfinite(p).all() and (sp <= RULES["vmax"] * (1 + rel)).all())
    rechecks = e.rechecks()
    # the first `steps` updates from the initial state (uniform start)
    load()
    e.sync()
    e.reset_timing()
    k0 = min(steps, T)
    t0 = time.perf_counter()
    e.step(k0)
    e.sync()
    w0 = time.perf_counter() - t0
    ms0, c0 = e.timing(E.K_PAIR)
    e.close()
    out = {"mode": E.MODE_NAMES[mode], "cells_per_axis": info.cells_per_axis, "block": info.block,
           "tile": info.tile, "fill": info.fill, "timed_updates": len(ts),
           "ms_per_step": 1e3 * wall / steps, "boid_steps_per_s": n * steps / wall,
           "pair_us_avg": 1e3 * pair_ms / max(pair_n, 1), "kernels_all_updates": kernels, "sane": ok,
           "fp64_rechecks": rechecks,
           "from_t0": {"updates": f"t = 0..{k0 - 1}", "ms_per_step": 1e3 * w0 / k0, "boid_steps_per_s": n * k0 / w0,
                       "pair_us_avg": 1e3 * ms0 / max(c0, 1)}}
    if samples:
        out["samples_t_ms_pairms"] = samples
    return out


def run_sweep(steps, base_steps, warmup, count=512, seed=3, block=0):
    """Batched flock table vs one engine per flock on the synthetic sweep."""
    from boidssimulatormpi_amd import sweep as S
    jobs = S.generated_sweep(count, seed)
    states = [S.initial_state(j) for j in jobs]
    total = sum(j.n for j in jobs)
    out = {"flocks": count, "boids": total, "seed": seed, "mode": "grid_mixed"}

    sw = S.Sweep(jobs, mode=E.MODE_GRID_MIXED, flags=E.FLAG_GRAPH | E.FLAG_TIMING, block=block)
    sw.set_initial_state(states)
    sw.step(warmup)
    sw.engine.sync()
    sw.engine.reset_timing()
    t0 = time.perf_counter()
    sw.step(steps)
    sw.engine.sync()
    wall = time.perf_counter() - t0
    ms_graph, _ = sw.engine.timing(E.K_STEP)
    p, v = sw.engine.state_f64()
    sane = bool(np.isfinite(p).all() and np.isfinite(v).all())
    sw.close()
    out["batched"] = {"steps": steps, "ms_per_step": 1e3 * wall / steps, "boid_steps_per_s": total * steps / wall,
                      "gpu_ms_per_step": ms_graph / steps, "sane": sane}
    log(f"[bench] sweep batched: {1e3 * wall / steps:.3f} ms/step")

    # The same kernels launched flock by flock.  per_flock / per_flock_graph
    # run the flocks one after another, as the reference runs a sweep (one job
    # per parameter set, tools/gen_jobscript.py:92-111); per_flock_streams
    # interleaves all flocks' updates on their engines' own streams.
    if base_steps <= 0:                 # batched only (no per-flock baselines)
        return out
    for label, flags, interleave in (("per_flock", 0, False), ("per_flock_graph", E.FLAG_GRAPH, False),
                                     ("per_flock_streams", 0, True)):
        engines = []
        for j, st in zip(jobs, states):
            e = E.Engine(j.params(E.MODE_GRID_MIXED, flags, block=block))
            e.set_state(xyzv6=st)
            engines.append(e)
        for e in engines:
            e.step(1)
            e.sync()
        t0 = time.perf_counter()
        if interleave:
            for _ in range(base_steps):
                for e in engines:
                    e.step(1)
            for e in engines:
                e.sync()
        else:
            for e in engines:
                e.step(base_steps)
                e.sync()
        wall = time.perf_counter() - t0
        for e in engines:
            e.close()
        out[label] = {"steps": base_steps, "ms_per_step": 1e3 * wall / base_steps,
                      "boid_steps_per_s": total * base_steps / wall}
        log(f"[bench] sweep {label}: {1e3 * wall / base_steps:.3f} ms/step")
    out["speedup_vs_per_flock"] = out["batched"]["boid_steps_per_s"] / out["per_flock"]["boid_steps_per_s"]
    out["speedup_vs_per_flock_graph"] = out["batched"]["boid_steps_per_s"] / out["per_flock_graph"]["boid_steps_per_s"]
    out["speedup_vs_per_flock_streams"] = out["batched"]["boid_steps_per_s"] / out["per_flock_streams"]["boid_steps_per_s"]
    return out


def cpu_info():
    """The host CPU as lscpu reports it, and the threads this process may use."""
    info = {"threads_available": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Vendor ID": "vendor", "CPU family": "family", "Model": "model_id",
                "Stepping": "stepping", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus", "CPU max MHz": "max_mhz",
                "L3": "l3", "Hypervisor vendor": "hypervisor"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in keys and keys[k] not in info:
                info[keys[k]] = v.strip()
    except Exception:
        pass
    return info


def _ref_update_seconds(exe, n, field, init, seed, threads, budget_s):
    """Seconds per update() of a reference engine (oracle/_ref/*, built from the
    reference's own sources by oracle/Makefile.ref): two runs of different
    length, so setup and init cancel.  Returns (seconds, updates timed)."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = RULES
    def run(t):
        args = [exe, os.devnull, str(n), str(t), repr(field), *(repr(x) for x in (*r["sep"], *r["ali"], *r["coh"])),
                repr(r["vmax"]), repr(r["vmin"]), init, str(seed), ""]
        t0 = time.perf_counter()
        subprocess.run(args, check=True, env=env, stdout=subprocess.DEVNULL, timeout=900)
        return time.perf_counter() - t0
    w1 = run(1)
    per = max(run(5) / 5, 1e-4)                          # for sizing only: start-up included, never an underestimate
    k = min(max(4, int(budget_s / per)), 100000)
    log(f"[bench]   {os.path.basename(exe)}, {threads} thread(s): {k} updates (~{k * per:.0f} s)")
    wk = run(1 + k)
    return (wk - w1) / k, k


def cpu_reference(which, n, field, seed, init, budget_s):
    """The reference's own engine on the same workload: ref_grid (boidsim_grid,
    src/boid_simulation_grid.cpp, with its own random_uniform init at `seed`)
    or ref_allpairs (boidsim, src/boid_simulation.cpp, reading `init` from a
    file), all host threads and one.  None when not built."""
    exe = os.path.join(ROOT, "oracle", "_ref", which)
    if not os.path.exists(exe):
        return None
    tmp = None
    if which == "ref_allpairs":
        import tempfile
        fd, tmp = tempfile.mkstemp(suffix=".init")
        os.close(fd)
        np.ascontiguousarray(init, "<f4").tofile(tmp)
        how = tmp
    else:
        how = "random_uniform"
    threads = len(os.sched_getaffinity(0))
    try:
        per, k = _ref_update_seconds(exe, n, field, how, seed, threads, budget_s)
        per1, k1 = _ref_update_seconds(exe, n, field, how, seed, 1, budget_s / 2)
    finally:
        if tmp:
            os.unlink(tmp)
    src = "src/boid_simulation_grid.cpp" if which == "ref_grid" else "src/boid_simulation.cpp"
    return {"value": n / per, "unit": "boid-steps/s", "cores": threads, "kind": "reference",
            "sample": f"{k} updates of oracle/_ref/{which} ({src} built from source), N={n}, {threads} threads",
            "value_1thread": n / per1, "sample_1thread": f"{k1} updates, 1 thread",
            "method": "(time of 1+k updates - time of 1 update) / k; init " +
                      ("the reference's random_uniform at the workload seed" if which == "ref_grid"
                       else "the workload's own initial state as an init file"),
            "cpu": cpu_info()}


def cpu_port(mid, t_mid, n, field, budget_s):
    """The oracle's FP64 cell-list update (the reference algorithm restated,
    OpenMP) on the same workload from its state at t = t_mid: full updates on
    all host threads until ~budget_s, and one thread on a sample of owners
    (phase A of every stride-th boid, scaled to all N, plus the timed cell-list
    build and phase B)."""
    from oracle import oracle as O
    P = O.params(field, RULES["sep"], RULES["ali"], RULES["coh"], RULES["vmax"], RULES["vmin"])
    p = np.ascontiguousarray(mid[0], np.float64).copy()
    v = np.ascontiguousarray(mid[1], np.float64).copy()
    threads = len(os.sched_getaffinity(0))
    # one thread first (the full updates below change the state)
    k, tb, ta, tp = O.time_step_sample(P, p, v, 1024, 1)
    want = max(1.0, budget_s / 2)
    stride = max(1, int(1024 * ta / want))
    k, tb, ta, tp = O.time_step_sample(P, p, v, stride, 1)
    per1 = tb + tp + ta * n / k
    steps, t0 = 0, time.perf_counter()
    while True:
        O.step(P, p, v, enumeration=O.ENUM_CELLS, nthreads=threads)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    wall = time.perf_counter() - t0
    return {"value": n * steps / wall, "unit": "boid-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} full updates from the t={t_mid} state, oracle FP64 cell list, {threads} threads, "
                      f"{wall:.1f} s",
            "value_1thread": n / per1,
            "sample_1thread": f"1 thread: phase A of every {stride}th owner ({k} owners, {ta:.1f} s) x {n / k:.1f}, "
                              f"+ cell-list build {tb:.2f} s + phase B {tp:.2f} s",
            "cpu": cpu_info()}


def traffic_from_profiles(mode_name, n):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the plain-copy calibration)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_pair_{mode_name}_n{n}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def valu_from_profiles(mode_name, n):
    """VALU issue fraction, active-lane fraction and instruction mix of the
    fused kernel from the committed rocprofv3 counter summary (every launch of
    a T = 1000 run)."""
    path = os.path.join(ROOT, "profiles", f"valu_pair_kernels_n{n}.txt")
    if not os.path.exists(path):
        return None
    keys = ("valu_issue", "active_lanes", "salu_per_valu", "lds_per_valu")
    got = {}
    with open(path) as f:
        for line in f:
            for key in keys:
                if line.startswith(f"{key}_{mode_name}:"):
                    got[key] = float(line.split(":", 1)[1])
    if "valu_issue" not in got or "active_lanes" not in got:
        return None
    out = {"issue_frac": got["valu_issue"], "active_lanes": got["active_lanes"],
           "source": os.path.relpath(path, ROOT), "launches": "every update of the T = 1000 run"}
    for key in ("salu_per_valu", "lds_per_valu"):
        if key in got:
            out[key] = got[key]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="flock1m", choices=sorted(WORKLOADS),
                    help="flock1m: the headline (N=2^20, F=4); flock16k: N=16384, F=1, the float grid mode "
                         "(also FP64 plain and mixed) vs the reference's own grid engine on the CPU; flock1k: "
                         "N=1024 all-pairs vs the reference's "
                         "own all-pairs engine; blobs16m: N=2^24, F=10, 64 clustered blobs")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed updates (default: the whole run); fewer than the run's length are stratified "
                         "over the run")
    ap.add_argument("--run-length", type=int, default=0,
                    help="updates in the run (default: the workload's configured T)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--from-t", type=int, default=0,
                    help="start the timed updates from the workload's state after this many updates (FP64 "
                         "engine, untimed; default 0: the run from its initial state)")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--field", type=float, default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--modes", default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="neighbour-kernel workgroup size (0: automatic)")
    ap.add_argument("--tune", action="store_true",
                    help="pick block/tile per mode with the autotuner (default: engine defaults, which are the "
                         "autotuner's choice for this workload)")
    ap.add_argument("--sweep-steps", type=int, default=None,
                    help="batched sweep steps (0: skip the sweep; default 100 with the headline workload)")
    ap.add_argument("--sweep-base-steps", type=int, default=10,
                    help="steps of the per-flock baselines (0: batched sweep only)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); the path's "
              "multi-GPU sharding is a design (spatial slabs with halo exchange), not a runnable mode.",
              file=sys.stderr)
        sys.exit(2)

    wl = WORKLOADS[a.workload]
    n = a.n if a.n is not None else wl["n"]
    F = a.field if a.field is not None else wl["field"]
    seed = a.seed if a.seed is not None else wl["seed"]
    T = a.run_length if a.run_length else wl["steps"]
    a.steps = a.steps if a.steps is not None else T
    a.modes = a.modes or wl["modes"]
    if a.sweep_steps is None:
        a.sweep_steps = 100 if wl["sweep"] else 0
    init = B.generate_init(wl["init"], n, F, RULES["vmax"], RULES["vmin"], seed=seed)
    by_name = {v: k for k, v in E.MODE_NAMES.items()}
    T_run = max(1, T - a.from_t)                     # updates from the start state to the run's end
    t_mid = T // 2
    mid = None
    if a.tune or (not a.no_cpu and wl["cpu"] == "port"):
        log(f"[bench] mid-run state (t={t_mid}) for tuning and the CPU sample")
        mid = mid_run_state(init, n, F, t_mid)
    res = {}
    start = init
    if a.from_t > 0:
        log(f"[bench] start state: t={a.from_t}")
        start = mid_run_state(init, n, F, a.from_t)
    ts = sample_times(T_run, a.steps)
    for name in a.modes.split(","):
        log(f"[bench] {name}: N={n} F={F} warmup={a.warmup} timed updates={a.steps} of a {T_run}-update run")
        params, tuning = tuned_params(by_name[name], mid, n, F, a.tune, a.block)
        res[name] = run_mode(by_name[name], start, n, F, T_run, a.steps, a.warmup, params)
        res[name]["tuning"] = tuning
        log(f"[bench] {name}: {res[name]['ms_per_step']:.3f} ms/step, {res[name]['boid_steps_per_s']:.3e} "
            f"boid-steps/s (from t=0 only: {res[name]['from_t0']['ms_per_step']:.3f} ms/step)")

    head = res.get(wl["head"]) or next(iter(res.values()))
    bytes_per = BYTES_PER_BOID_STEP // 2 if head["mode"] in ("allpairs_f32ref", "grid_f32ref") else BYTES_PER_BOID_STEP
    achieved = bytes_per * n / (head["pair_us_avg"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(head["mode"], n),
                "kernel": f"fused neighbour+integrate ({head['mode']})", "bytes_per_boid_step": bytes_per,
                "kernel_us_avg": head["pair_us_avg"], "over": "the timed updates (HIP events on the engine's stream)",
                "valu": valu_from_profiles(head["mode"], n)}
    sweep = None
    if a.sweep_steps > 0:
        log(f"[bench] sweep: 512 flocks, {a.sweep_steps} batched steps, {a.sweep_base_steps} per-flock steps")
        sweep = run_sweep(a.sweep_steps, a.sweep_base_steps, 5, block=a.block)
    cpu = None
    if not a.no_cpu and wl["cpu"]:        # blobs16m: one oracle step alone exceeds the sample budget
        log(f"[bench] cpu baseline (~{a.cpu_seconds:.0f} s all threads + ~{a.cpu_seconds / 2:.0f} s one thread)")
        cpu = (cpu_port(mid, t_mid, n, F, a.cpu_seconds) if wl["cpu"] == "port"
               else cpu_reference(wl["cpu"], n, F, seed, init, a.cpu_seconds))
    metric = ("boid-steps/sec at N=2^20: FP64 plain vs mixed precision; % HBM roofline" if a.workload == "flock1m"
              else f"boid-steps/sec, {a.workload} (N={n}, {head['mode']})")
    if a.steps >= T_run:
        timed = f"the whole {a.steps}-update run from t = {a.from_t}, consecutive updates"
    else:
        at = ", ".join(str(t + a.from_t) for t in ts) if len(ts) <= 3 else \
            f"{ts[0] + a.from_t}, {ts[1] + a.from_t}, ..., {ts[-1] + a.from_t}"
        timed = (f"{a.steps} single updates stratified over the {T_run}-update run from t = {a.from_t} "
                 f"(t = {at}: midpoints of {a.steps} equal segments), each bracketed by device syncs, "
                 f"the rest untimed")
    dtype = {"grid_mixed": "f32+f64", "allpairs_f32ref": "f32", "grid_f32ref": "f32"}.get(head["mode"], "f64")
    out = {
        "metric": metric,
        "value": head["boid_steps_per_s"],
        "unit": "boid-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": f"synthetic ({wl['what']}, seed {seed})",
        "config": {"workload": f"boid update, N={n}, F={F}, {wl['what']}, |v| in [0.001,0.005], radii "
                               f"0.02/0.05/0.08, view 270 deg, weights 0.002/0.06/0.008, seed {seed}, "
                               f"a {T}-update run",
                   "timed": timed, "name": a.workload, "n": n, "field": F, "run_length": T,
                   "timed_from_t": a.from_t, "headline_mode": head["mode"], "parallelism": "single GPU"},
        "value_from_t0": head["from_t0"]["boid_steps_per_s"],
        "modes": res,
        "mixed_over_fp64": (res["grid_mixed"]["boid_steps_per_s"] / res["grid_f64"]["boid_steps_per_s"]
                            if "grid_mixed" in res and "grid_f64" in res else None),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
