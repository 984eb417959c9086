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
    assert set(bench.WORKLOADS) == {"flock1m", "flock16k", "blobs16m"}
    for name, w in bench.WORKLOADS.items():
        assert w["head"] in w["modes"].split(","), name
        assert w["cpu"] in ("port", "reference", None), name
    h = bench.WORKLOADS["flock1m"]
    assert (h["n"], h["field"], h["seed"], h["steps"], h["head"]) == (1 << 20, 4.0, 2, 1000, "grid_mixed")
