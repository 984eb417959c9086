"""Batched sweeps on the GPU: a flock table stepped by one launch per kernel
gives, flock by flock, exactly what a single-flock engine gives (which the
parity tests tie to the oracle); graph-replayed updates equal eager ones;
per-flock metrics equal single-engine metrics; and the sweep front-end
writes the same .ptcl bytes as one boidsim_grid run per job
(tools/gen_jobscript.py:92-111 runs the sweep that way)."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

from boidssimulatormpi_amd import engine as E
from boidssimulatormpi_amd import sweep as S

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
BIN = os.path.join(ROOT, "boidssimulatormpi_amd", "bin")

SWEEP_INI = """[Global]
POPULATION = 2048
FIELD_SIZE = 0.5
INIT = random_uniform
RANDOM_SEED = 12345

[Separation]
SIGHT_DISTANCE = 0.01
SIGHT_ANGLE= 0.5
FORCE_COEFFICIENT = 0.002

[Alignment]
SIGHT_DISTANCE = 0.05
SIGHT_ANGLE= 0.3333333333333333
FORCE_COEFFICIENT = 0.06

[Cohesion]
SIGHT_DISTANCE = 0.05
SIGHT_ANGLE= 0.5
FORCE_COEFFICIENT = 0.008

[Velocity]
MIN = {vmin}
MAX = {vmax}
"""


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_jobs(count, seed, nmax):
    """Flocks of the synthetic sweep, populations capped for test time."""
    jobs = S.generated_sweep(count, seed)
    out = []
    for j in jobs:
        n = min(j.n, nmax)
        out.append(S.Job(name=j.name, n=n, field=(n / 16384.0) ** (1 / 3), sep=j.sep, ali=j.ali, coh=j.coh,
                         vmax=j.vmax, vmin=j.vmin, init="isotropic", seed=j.seed, time_step=j.time_step))
    return out


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_batch_equals_single_engines(mode):
    jobs = small_jobs(7, 11, 9000)
    flags = E.FLAG_FORCES | E.FLAG_DIGEST
    say(f"batch of {len(jobs)} flocks, mode {E.MODE_NAMES[mode]}, n = {[j.n for j in jobs]}")
    sw = S.Sweep(jobs, mode=mode, flags=flags)
    assert sw.engine.flock_count() == len(jobs)
    for k, j in enumerate(jobs):
        fi = sw.engine.flock_info(k)
        assert fi.boid_base == sw.base[k] and fi.n == j.n and fi.cells_per_axis >= 3
    states = [S.initial_state(j) for j in jobs]
    sw.set_initial_state(states)
    singles = []
    for j, st in zip(jobs, states):
        e = E.Engine(j.params(mode, flags))
        e.set_state(xyzv6=st)
        singles.append(e)
    assert sw.engine.info().mode == mode and all(e.info().mode == mode for e in singles)
    for t in range(1, 4):
        sw.step(1)
        for e in singles:
            e.step(1)
        pb, vb = sw.states_f64()
        fb = sw.split(sw.engine.forces_f64())
        db = sw.split(sw.engine.digest())
        for k, e in enumerate(singles):
            p, v = e.state_f64()
            assert np.array_equal(p, pb[k]) and np.array_equal(v, vb[k]), (t, k)
            assert np.array_equal(e.forces_f64(), fb[k]), (t, k)
            assert np.array_equal(e.digest()[:, 0], db[k][:, 0]), (t, k)   # neighbour counts per rule
    for e in singles:
        e.close()
    sw.close()


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_graph_replays_match_eager(mode):
    n, F = 30000, 1.2
    st = E.generate_init(E.INIT_ISOTROPIC, n, F, 0.005, 0.001, seed=5)
    say(f"graph vs eager n={n} mode {E.MODE_NAMES[mode]}")
    g = E.Engine(E.make_params(n, F, mode=mode, flags=E.FLAG_GRAPH | E.FLAG_TIMING))
    e = E.Engine(E.make_params(n, F, mode=mode))
    g.set_state(xyzv6=st)
    e.set_state(xyzv6=st)
    for k in (1, 1, 3, 5):                     # host reads and syncs between replays
        g.step(k)
        e.step(k)
        pg, vg = g.state_f64()
        pe, ve = e.state_f64()
        assert np.array_equal(pg, pe) and np.array_equal(vg, ve), k
    ms, launches = g.timing(E.K_STEP)
    assert launches == 4 and ms > 0.0
    assert g.info().steps == 10
    g.close()
    e.close()


def test_batch_graph_and_metrics():
    jobs = small_jobs(5, 13, 6000)
    st = [S.initial_state(j) for j in jobs]
    a = S.Sweep(jobs, mode=E.MODE_GRID_MIXED, flags=E.FLAG_GRAPH)
    b = S.Sweep(jobs, mode=E.MODE_GRID_MIXED, flags=0)
    a.set_initial_state(st)
    b.set_initial_state(st)
    a.step(4)
    b.step(4)
    pa, va = a.states_f64()
    pb, vb = b.states_f64()
    for k in range(len(jobs)):
        assert np.array_equal(pa[k], pb[k]) and np.array_equal(va[k], vb[k])
    with pytest.raises(E.EngineError):
        a.engine.metrics(0.05)                  # whole-batch metrics would mix flocks
    for k, j in enumerate(jobs):
        e = E.Engine(j.params(E.MODE_GRID_MIXED))
        e.set_state(pos=pa[k], vel=va[k])
        assert a.engine.metrics(0.05, 2, flock=k) == e.metrics(0.05, 2)
        e.close()
    a.close()
    b.close()


def test_sweep_front_end_matches_boidsim_grid(tmp_path):
    settings = []
    for vmin, vmax in ((0.001, 0.003), (0.002, 0.004)):
        p = tmp_path / f"vmin{vmin:.4f}_vmax{vmax:.4f}.ini"
        p.write_text(SWEEP_INI.format(vmin=f"{vmin:.4f}", vmax=f"{vmax:.4f}"))
        settings.append(str(p))
    T = 6
    jobs = S.jobs_from_settings(settings, [2048, 5000], 16384.0, T)
    say("sweep front-end:", [j.name for j in jobs])
    sw = S.run(jobs, str(tmp_path / "sweep"), mode=E.MODE_GRID_F64, log=None)
    sw.close()
    exe = os.path.join(BIN, "boidsim_grid")
    assert os.path.exists(exe), "boidsim_grid missing: run __graft_entry__.build()"
    for j in jobs:
        out = tmp_path / (j.name + ".ptcl")
        r = subprocess.run([exe, "-s", j.setting, "-T", str(T), "-N", str(j.n), "-F", repr(j.field), "-o", str(out)],
                           capture_output=True, text=True, timeout=300,
                           env={**os.environ, "BOIDS_ENGINE_MODE": "grid_f64"})
        assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-1000:]
        a = hashlib.sha256(open(out, "rb").read()).digest()
        b = hashlib.sha256(open(tmp_path / "sweep" / (j.name + ".ptcl"), "rb").read()).digest()
        assert a == b, j.name
