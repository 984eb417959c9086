"""Parameter sweeps as one batched engine (bs_create_batch).

The reference runs a sweep as many independent jobs: gen.py writes one .ini
per parameter set (settings/parameter_search_vel_2/gen.py:32-42) and
gen_jobscript.py turns (setting, population) pairs into one boidsim job each,
with the field sized from a density (tools/gen_jobscript.py:20-21, 97-111):

    name  = "{setting stem}_N{n:07d}_T{t}"
    field = (n / density) ** (1/3)
    ./boidsim -s setting -T t -N n -F field -o name.ptcl

Here the same job list becomes one flock table: every job is a flock of one
engine, one launch per kernel steps them all, and each job still gets the
.ptcl file its boidsim_grid run would write (same header, frames and bytes
when the mode is grid_f64).

Command line (gen_jobscript's arguments, then run instead of submit):
    python -m boidssimulatormpi_amd.sweep -s a.ini b.ini -n "2**10" "[2048,4096]" \
        -d 16384 -t 100 -o outdir [--mode grid_f64|grid_mixed] [--no-frames]
"""
from __future__ import annotations

import argparse
import ast
import math
import os
import re
import sys
import time
from dataclasses import dataclass, field as dfield

import numpy as np

from . import engine as E
from .ptcl import HEADER_SIZE, Header

FPS = 30


@dataclass
class Job:
    """One flock of a sweep: the settings one boidsim run would use."""
    name: str
    n: int
    field: float
    sep: tuple
    ali: tuple
    coh: tuple
    vmax: float
    vmin: float
    init: str = "random_uniform"        # "test", "random_uniform", "isotropic" or an init file
    seed: int = 0
    time_step: int = 0
    setting: str = ""
    extra: dict = dfield(default_factory=dict)

    def params(self, mode=E.MODE_GRID_MIXED, flags=0, device=0, block=0) -> E.Params:
        return E.make_params(self.n, self.field, self.sep, self.ali, self.coh, self.vmax, self.vmin,
                             mode=mode, flags=flags, device=device, block=block)


# ----------------------------------------------------------------------------
# .ini settings, read as the drop-in driver reads them (csrc/args.cpp; the
# reference: src/args.cpp:78-153 over Boost.PropertyTree's INI parser)
# ----------------------------------------------------------------------------
_INT = re.compile(r"^[+-]?\d+$")
_FLOAT = re.compile(r"^[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)$")


def read_ini(path) -> dict:
    """{"Section.KEY": "value"}: [Section] headers, KEY = value lines, lines
    starting with ';' or '#' ignored, surrounding blanks trimmed."""
    kv = {}
    section = ""
    with open(path) as f:
        for line in f:
            t = line.strip()
            if not t or t[0] in ";#":
                continue
            if t[0] == "[":
                e = t.find("]")
                section = (t[1:e] if e >= 0 else t[1:]).strip()
                continue
            eq = t.find("=")
            if eq < 0:
                continue
            kv[section + "." + t[:eq].strip()] = t[eq + 1:].strip()
    return kv


def _num(kv, key, default, kind):
    v = kv.get(key)
    if v is None:
        return default
    if kind is int:
        return int(v) if _INT.match(v) else default
    return float(v) if _FLOAT.match(v) else default


def read_settings(path) -> dict:
    """The Args fields a settings file sets (absent or malformed keys keep
    the defaults: 0 and an empty INIT)."""
    kv = read_ini(path)
    return {
        "time_step": _num(kv, "Global.TIME_STEP", 0, int),
        "population": _num(kv, "Global.POPULATION", 0, int),
        "field_size": _num(kv, "Global.FIELD_SIZE", 0.0, float),
        "init": kv.get("Global.INIT", ""),
        "seed": _num(kv, "Global.RANDOM_SEED", 0, int),
        "sep": (_num(kv, "Separation.SIGHT_DISTANCE", 0.0, float), _num(kv, "Separation.SIGHT_ANGLE", 0.0, float),
                _num(kv, "Separation.FORCE_COEFFICIENT", 0.0, float)),
        "ali": (_num(kv, "Alignment.SIGHT_DISTANCE", 0.0, float), _num(kv, "Alignment.SIGHT_ANGLE", 0.0, float),
                _num(kv, "Alignment.FORCE_COEFFICIENT", 0.0, float)),
        "coh": (_num(kv, "Cohesion.SIGHT_DISTANCE", 0.0, float), _num(kv, "Cohesion.SIGHT_ANGLE", 0.0, float),
                _num(kv, "Cohesion.FORCE_COEFFICIENT", 0.0, float)),
        "vmax": _num(kv, "Velocity.MAX", 0.0, float),
        "vmin": _num(kv, "Velocity.MIN", 0.0, float),
    }


def job_name(setting, n, t):
    """tools/gen_jobscript.py:20-21."""
    return "{}_N{}_T{}".format(os.path.splitext(os.path.basename(setting))[0], str(n).zfill(7), t)


def populations(exprs) -> list:
    """gen_jobscript's -n values: integers, a**b powers, lists and
    range(...) comprehension-free lists, flattened (tools/gen_jobscript.py:9-16,93)."""
    out = []

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, int):
            return node.value
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Pow, ast.Mult, ast.Add, ast.Sub)):
            a, b = ev(node.left), ev(node.right)
            return {ast.Pow: lambda: a ** b, ast.Mult: lambda: a * b, ast.Add: lambda: a + b,
                    ast.Sub: lambda: a - b}[type(node.op)]()
        if isinstance(node, (ast.List, ast.Tuple)):
            return [ev(e) for e in node.elts]
        if isinstance(node, ast.Call) and getattr(node.func, "id", None) == "range" and not node.keywords:
            return list(range(*[ev(a) for a in node.args]))
        raise ValueError("unsupported population expression")

    def flat(v):
        if isinstance(v, list):
            for x in v:
                flat(x)
        else:
            out.append(int(v))

    for s in exprs:
        flat(ev(ast.parse(s, mode="eval")))
    return out


def jobs_from_settings(settings, pops, density, time_step) -> list:
    """(setting, population) jobs as gen_jobscript builds them
    (tools/gen_jobscript.py:92-111): N and T from the command line, the
    field from the density, everything else from the setting file."""
    jobs = []
    for s in settings:
        st = read_settings(s)
        for n in pops:
            jobs.append(Job(name=job_name(s, n, time_step), n=n, field=(n / density) ** (1.0 / 3.0),
                            sep=st["sep"], ali=st["ali"], coh=st["coh"], vmax=st["vmax"], vmin=st["vmin"],
                            init=st["init"], seed=st["seed"], time_step=time_step, setting=s))
    return jobs


def generated_sweep(count=512, seed=3, density=16384.0, time_step=2000,
                    base=((0.02, 0.75, 0.002), (0.05, 0.75, 0.06), (0.08, 0.75, 0.008))) -> list:
    """The synthetic many-flock sweep: N log-uniform in [2^10, 2^16],
    field (N/density)^(1/3), vmax ~ U[0.002, 0.01], vmin = vmax/5, all radii
    scaled by s ~ U[0.7, 1.3]; isotropic init seeded per flock.  Draws come
    from numpy's PCG64 stream of `seed` in the order u_N, vmax, s per flock."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        u, vmax, s = rng.random(), rng.uniform(0.002, 0.01), rng.uniform(0.7, 1.3)
        n = int(math.floor(2.0 ** (10.0 + 6.0 * u)))
        sep, ali, coh = [(r * s, a, w) for (r, a, w) in base]
        jobs.append(Job(name=f"flock{k:04d}", n=n, field=(n / density) ** (1.0 / 3.0), sep=sep, ali=ali, coh=coh,
                        vmax=vmax, vmin=vmax / 5.0, init="isotropic", seed=seed * 100000 + k,
                        time_step=time_step))
    return jobs


_KINDS = {"test": E.INIT_TEST, "random_uniform": E.INIT_RANDOM_UNIFORM, "isotropic": E.INIT_ISOTROPIC,
          "clustered": E.INIT_CLUSTERED}


def initial_state(job: Job) -> np.ndarray:
    """N x 6 float32 init of one job, as the grid engine makes it
    (srand48(seed) first, src/boid_simulation_grid.cpp:98), or read from an
    init file (src/boid_simulation.cpp:113-136)."""
    if job.init in _KINDS:
        return E.generate_init(_KINDS[job.init], job.n, job.field, job.vmax, job.vmin, seed=job.seed)
    import ctypes
    st = np.zeros((job.n, 6), np.float32)
    rc = E.lib().bs_read_init_file(job.init.encode(), job.n, E._p(st, ctypes.c_float))
    if rc != 0:
        raise E.EngineError(E.lib().bs_last_error(None).decode())
    return st


class Sweep:
    """A batched engine over a job list, plus per-job views of its arrays."""

    def __init__(self, jobs, mode=E.MODE_GRID_MIXED, flags=E.FLAG_GRAPH, device=0, block=0):
        if not jobs:
            raise ValueError("empty sweep")
        self.jobs = list(jobs)
        self.engine = E.Engine([j.params(mode, flags, device, block) for j in self.jobs])
        self.base = np.cumsum([0] + [j.n for j in self.jobs])

    def set_initial_state(self, states=None):
        states = states if states is not None else [initial_state(j) for j in self.jobs]
        self.engine.set_state(xyzv6=np.concatenate(states, axis=0))

    def step(self, nsteps=1):
        self.engine.step(nsteps)

    def split(self, a):
        """Per-job slices of an id-ordered array."""
        return [a[self.base[k]:self.base[k + 1]] for k in range(len(self.jobs))]

    def states_f32(self):
        p, v = self.engine.state_f32()
        return self.split(p), self.split(v)

    def states_f64(self):
        p, v = self.engine.state_f64()
        return self.split(p), self.split(v)

    def close(self):
        self.engine.close()


def run(jobs, outdir, mode=E.MODE_GRID_MIXED, frames=True, log=sys.stdout) -> Sweep:
    """Runs every job for its time_step updates in one batched engine and,
    with frames, writes <outdir>/<name>.ptcl per job exactly as the
    reference driver does (header, then the positions of t = 0 .. T-1,
    src/main.cpp:187-292).  All jobs must share one time_step."""
    steps = {j.time_step for j in jobs}
    if len(steps) != 1:
        raise ValueError("all jobs of one batch need the same time step count")
    T = steps.pop()
    sw = Sweep(jobs, mode=mode)
    sw.set_initial_state()
    files = []
    if frames:
        os.makedirs(outdir, exist_ok=True)
        for j in jobs:
            f = open(os.path.join(outdir, j.name + ".ptcl"), "wb")
            f.write(Header(j.n, T, j.field, t_0=0, fps=FPS).pack())
            files.append(f)
    t0 = time.perf_counter()
    for t in range(T):
        if frames:
            pos, _ = sw.states_f32()
            for f, p in zip(files, pos):
                f.write(np.ascontiguousarray(p, np.float32).tobytes())
        sw.step(1)
        if log and (t % max(1, T // 10) == 0 or t == T - 1):
            print(f"  {t}/{T}", file=log, flush=True)
    sw.engine.sync()
    for f in files:
        f.close()
    if log:
        print(f"sweep of {len(jobs)} flocks ({int(sw.base[-1])} boids), {T} steps: "
              f"{time.perf_counter() - t0:.2f} s", file=log, flush=True)
    return sw


def _main(argv=None):
    ap = argparse.ArgumentParser(description="Run a parameter sweep as one batched engine.")
    ap.add_argument("-s", dest="settings", nargs="+", required=True, help="setting files (.ini)")
    ap.add_argument("-n", dest="populations", nargs="+", required=True,
                    help="populations: integers, a**b, lists or range(...)")
    ap.add_argument("-d", dest="density", type=float, required=True, help="boids per unit volume")
    ap.add_argument("-t", dest="time_step", type=int, required=True, help="time steps")
    ap.add_argument("-o", dest="output_dir", default=".", help="directory for the .ptcl files")
    ap.add_argument("--mode", choices=["grid_f64", "grid_mixed"], default="grid_mixed")
    ap.add_argument("--no-frames", action="store_true", help="run without writing .ptcl files")
    a = ap.parse_args(argv)
    jobs = jobs_from_settings(a.settings, populations(a.populations), a.density, a.time_step)
    mode = E.MODE_GRID_F64 if a.mode == "grid_f64" else E.MODE_GRID_MIXED
    sw = run(jobs, a.output_dir, mode=mode, frames=not a.no_frames)
    sw.close()
    return 0


if __name__ == "__main__":
    sys.exit(_main())
