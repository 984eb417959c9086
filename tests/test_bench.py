"""CPU checks of bench.py's contract: it measures one GPU only, reads the
figures it reports from the committed rocprofv3 summaries, and its workload
table is consistent.  (The timed runs themselves need a GPU.)"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("mode", ["grid_mixed", "grid_f64"])
def test_profile_figures_are_read(mode):
    t = bench.traffic_from_profiles(mode, 1 << 20)
    assert t is not None and t > bench.BYTES_PER_BOID_STEP * (1 << 20)   # measured bytes exceed the algorithmic ones
    v = bench.valu_from_profiles(mode, 1 << 20)
    assert v is not None and 0 < v["issue_frac"] <= 1 and 0 < v["active_lanes"] <= 1
    assert os.path.exists(os.path.join(ROOT, v["source"]))
    assert bench.traffic_from_profiles(mode, 12345) is None and bench.valu_from_profiles(mode, 12345) is None


def test_workloads():
    assert set(bench.WORKLOADS) == {"flock1m", "flock16k", "flock1k", "blobs16m"}
    for name, w in bench.WORKLOADS.items():
        assert w["head"] in w["modes"].split(","), name
        assert w["cpu"] in ("port", "ref_grid", "ref_allpairs", None), name
    h = bench.WORKLOADS["flock1m"]
    assert (h["n"], h["field"], h["seed"], h["steps"], h["head"]) == (1 << 20, 4.0, 2, 1000, "grid_mixed")
    c0 = bench.WORKLOADS["flock1k"]
    assert (c0["n"], c0["field"], c0["seed"], c0["steps"], c0["head"]) == (1024, 1.0, 0, 100, "allpairs_f64")


def test_stratified_sample_times():
    # K < T: the midpoints of K equal segments, strictly increasing, inside the run
    ts = bench.sample_times(1000, 20)
    assert ts == [25 + 50 * k for k in range(20)]
    for T, K in ((1000, 7), (8000, 20), (500, 499), (3, 2)):
        ts = bench.sample_times(T, K)
        assert len(ts) == K and all(a < b for a, b in zip(ts, ts[1:])) and 0 <= ts[0] and ts[-1] < T
        # segments of at least one update hold exactly their sample
        if T >= K:
            assert all(k * T / K - 1 < t < (k + 1) * T / K for k, t in enumerate(ts))
        if T >= 2 * K:
            assert all(k * T / K <= t < (k + 1) * T / K for k, t in enumerate(ts))
    # K >= T: the whole run, consecutive
    assert bench.sample_times(100, 100) == list(range(100))
    assert bench.sample_times(100, 150) == list(range(150))


def test_cpu_port_sampler_estimates_a_full_step():
    # the one-thread figure comes from phase A of a sample of owners: on a
    # small flock the scaled estimate matches a timed full update
    import time

    import numpy as np

    import boidssimulatormpi_amd as B
    from boidssimulatormpi_amd import engine as E
    from oracle import oracle as O
    n, F = 32768, (32768 / 16384.0) ** (1.0 / 3.0)
    st = B.generate_init(E.INIT_ISOTROPIC, n, F, 0.005, 0.001, seed=3)
    P = O.params(F)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    p0, v0 = p.copy(), v.copy()
    # interleaved repeats, the fastest of each: timings taken while other
    # work shares the host stay comparable
    ests, fulls = [], []
    for _ in range(3):
        k, tb, ta, tp = O.time_step_sample(P, p, v, 8, 1)
        assert k == n // 8 and np.array_equal(p, p0) and np.array_equal(v, v0)   # state untouched
        ests.append(tb + tp + ta * n / k)
        fulls.append(_timed_step(O, P, p0, v0))
    est, best = min(ests), min(fulls)
    # idle host: within ~7 %; the bound catches a mis-scaled sample (the
    # stride is 8) while tolerating a loaded host
    assert 0.5 < est / best < 2.0, (ests, fulls)


def _timed_step(O, P, p0, v0):
    import time
    p, v = p0.copy(), v0.copy()
    t0 = time.perf_counter()
    O.step(P, p, v, enumeration=O.ENUM_CELLS, nthreads=1)
    return time.perf_counter() - t0
