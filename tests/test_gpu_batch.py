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
            assert np.array_equal(e.digest(), db[k]), (t, k)   # neighbour sets (flock-local ids)
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


@pytest.mark.parametrize("mode", [E.MODE_GRID_MIXED, E.MODE_GRID_F64])
def test_cfg3_full_sweep_against_oracle(mode):
    # The configured sweep itself: 512 flocks, N log-uniform in [2^10, 2^16]
    # (8.7 M boids), per-flock fields, vmax ~ U[0.002, 0.01], vmin = vmax / 5,
    # radii x U[0.7, 1.3], seed 3 (settings/parameter_search_vel_2/gen.py:32-42
    # and tools/gen_jobscript.py:37-65 are the reference's form of a sweep:
    # one job per parameter set).  One batched engine steps every flock; the
    # extreme flocks (smallest and largest N, smallest and largest radius
    # scale and vmax) are checked against the FP64 oracle on their own rules
    # every step: neighbour sets exact, per-step error <= 1e-12 (FP64) or
    # 1e-6 (mixed) of the field and of the flock's vmax.  Every flock stays in
    # its field with vmin <= |v| <= vmax.
    from oracle import oracle as O
    jobs = S.generated_sweep(512, 3)
    scale = [j.coh[0] / 0.08 for j in jobs]
    picks = sorted({int(np.argmin([j.n for j in jobs])), int(np.argmax([j.n for j in jobs])),
                    int(np.argmin(scale)), int(np.argmax(scale)),
                    int(np.argmin([j.vmax for j in jobs])), int(np.argmax([j.vmax for j in jobs]))})
    say(f"cfg3 {E.MODE_NAMES[mode]}: {sum(j.n for j in jobs)} boids in {len(jobs)} flocks; oracle on flocks {picks}: "
        + ", ".join(f"n={jobs[k].n} s={scale[k]:.3f} vmax={jobs[k].vmax:.4f}" for k in picks))
    e = E.Engine([j.params(mode, E.FLAG_DIGEST) for j in jobs])
    assert e.flock_count() == 512
    e.set_state(xyzv6=np.concatenate([S.initial_state(j) for j in jobs]))
    bases = [e.flock_info(k).boid_base for k in range(512)]
    tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
    gp, gv = e.state_f64()
    worst = [0.0, 0.0]
    for t in range(3):
        want = {}
        for k in picks:
            j, b = jobs[k], bases[k]
            P = O.params(j.field, j.sep, j.ali, j.coh, j.vmax, j.vmin)
            p, v = gp[b:b + j.n].copy(), gv[b:b + j.n].copy()
            _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
            want[k] = (p, v, d)
        e.step(1)
        gd = e.digest()
        gp, gv = e.state_f64()
        for k in picks:
            j, b = jobs[k], bases[k]
            p, v, d = want[k]
            bad = np.nonzero((gd[b:b + j.n] != d).any(1))[0]   # a batch's digests hash flock-local ids
            assert bad.size == 0, (t, k, f"{bad.size} boids with different neighbour sets")
            dp = gp[b:b + j.n] - p
            dp -= j.field * np.round(dp / j.field)
            ep = float(np.abs(dp).max() / j.field)
            ev = float(np.abs(gv[b:b + j.n] - v).max() / j.vmax)
            worst = [max(worst[0], ep), max(worst[1], ev)]
            assert ep <= tol and ev <= tol, (t, k, ep, ev)
        for k, j in enumerate(jobs):
            b = bases[k]
            p, v = gp[b:b + j.n], gv[b:b + j.n]
            s = np.linalg.norm(v, axis=1)
            assert np.isfinite(p).all() and (p >= 0).all() and (p <= j.field).all(), (t, k)
            assert (s >= j.vmin * (1 - 1e-12)).all() and (s <= j.vmax * (1 + 1e-12)).all(), (t, k)
        say(f"  step {t}: ok")
    say(f"cfg3 {E.MODE_NAMES[mode]}: max per-step error {worst[0]:.3g} F, {worst[1]:.3g} vmax; "
        f"rechecks {e.rechecks()}")
    e.close()
