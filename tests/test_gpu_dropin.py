"""The drop-in executables on the GPU: boidsim reproduces the reference
boidsim's .ptcl output byte for byte (settings/test.ini golden digests), the
split/resume invariant of test/compare_split.sh holds with the executables,
--force-data-output writes the reference's 12-float records, boidsim_grid
reproduces the reference boidsim_grid's output byte for byte, and its FP64
mode tracks the FP64 oracle."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from boidssimulatormpi_amd import ptcl
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
BIN = os.path.join(ROOT, "boidssimulatormpi_amd", "bin")
TESTINI = os.path.join(ROOT, "tests", "golden", "settings_test.ini")
TESTINI_RULES = ((0.1, 0.9, 0.0001), (0.1, 0.7, 0.1), (0.1, 0.8, 0.1), 0.003, 0.001)


def run(exe, *args, cwd, env=None):
    path = os.path.join(BIN, exe)
    assert os.path.exists(path), f"{path} missing: run __graft_entry__.build()"
    print("run", exe, *args, flush=True)
    r = subprocess.run([path, *args], cwd=cwd, capture_output=True, text=True, timeout=300,
                       env={**os.environ, **(env or {})})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).digest()


def test_boidsim_matches_reference_output(golden_dir, tmp_path):
    g = np.load(os.path.join(golden_dir, "ref_allpairs_testini_n1000.npz"))
    out = run("boidsim", "-s", TESTINI, "-o", "a.ptcl", "--velocity-output", "av.ptcl", cwd=tmp_path)
    assert "simulation finished." in out and "  29/30" in out
    assert sha(tmp_path / "a.ptcl") == g["ptcl_sha256"].tobytes()
    assert sha(tmp_path / "av.ptcl") == g["vptcl_sha256"].tobytes()


def test_boidsim_split_resume_equals_continuous(tmp_path):
    run("boidsim", "-s", TESTINI, "-T", "30", "-o", "full.ptcl", cwd=tmp_path)
    run("boidsim", "-s", TESTINI, "-T", "15", "-o", "a.ptcl", "--velocity-output", "av.ptcl", cwd=tmp_path)
    ptcl.generate_initialization_file(tmp_path / "a.ptcl", tmp_path / "av.ptcl", 14, tmp_path / "init.bin")
    run("boidsim", "-s", TESTINI, "-T", "16", "-o", "b.ptcl", "--initialization", str(tmp_path / "init.bin"),
        cwd=tmp_path)
    ptcl.split(tmp_path / "b.ptcl", tmp_path / "tmp.ptcl", 1, 15)
    ptcl.merge([tmp_path / "a.ptcl", tmp_path / "tmp.ptcl"], tmp_path / "merged.ptcl")
    assert sha(tmp_path / "merged.ptcl") == sha(tmp_path / "full.ptcl")


def test_boidsim_force_output(tmp_path):
    run("boidsim", "-s", TESTINI, "-T", "3", "-N", "200", "-o", "f.ptcl", "--force-data-output", cwd=tmp_path)
    f = ptcl.PtclFile(tmp_path / "f.ptcl", force=True)
    assert f.check_metadata() and f.valid_steps == 3
    fr = f.frames()
    P = O.params(1.0, *TESTINI_RULES)
    p, v = O.init_reference("test", 200, 1.0, 0.003, 0.001)
    w = np.float32([0.1] * 3 + [0.0001] * 3 + [0.1] * 3)
    assert np.array_equal(fr[0, :, :3], p) and not fr[0, :, 3:].any()    # zeros before the first update
    for t in (1, 2):
        frc, _ = O.step(P, p, v, forces=True)
        assert np.array_equal(fr[t, :, :3], p) and np.array_equal(fr[t, :, 3:], frc * w), t


def test_boidsim_grid_matches_reference_output(golden_dir, tmp_path):
    # boidsim_grid runs the float grid mode: the reference boidsim_grid's
    # .ptcl and velocity files for settings/test.ini, byte for byte (its own
    # init: srand48(12345) then the "test" kind; GRID_NUM 4)
    g = np.load(os.path.join(golden_dir, "ref_grid_testini_n1000.npz"))
    out = run("boidsim_grid", "-s", TESTINI, "-o", "g.ptcl", "--velocity-output", "gv.ptcl", cwd=tmp_path)
    assert "grid num:" in out and "simulation finished." in out
    assert sha(tmp_path / "g.ptcl") == g["ptcl_sha256"].tobytes()
    assert sha(tmp_path / "gv.ptcl") == g["vptcl_sha256"].tobytes()


def test_boidsim_grid_split_resume_equals_continuous(tmp_path):
    # test/compare_split.sh with the grid executable (the grid is rebuilt
    # after a file init)
    run("boidsim_grid", "-s", TESTINI, "-T", "30", "-o", "full.ptcl", cwd=tmp_path)
    run("boidsim_grid", "-s", TESTINI, "-T", "15", "-o", "a.ptcl", "--velocity-output", "av.ptcl", cwd=tmp_path)
    ptcl.generate_initialization_file(tmp_path / "a.ptcl", tmp_path / "av.ptcl", 14, tmp_path / "init.bin")
    run("boidsim_grid", "-s", TESTINI, "-T", "16", "-o", "b.ptcl", "--initialization", str(tmp_path / "init.bin"),
        cwd=tmp_path)
    ptcl.split(tmp_path / "b.ptcl", tmp_path / "tmp.ptcl", 1, 15)
    ptcl.merge([tmp_path / "a.ptcl", tmp_path / "tmp.ptcl"], tmp_path / "merged.ptcl")
    assert sha(tmp_path / "merged.ptcl") == sha(tmp_path / "full.ptcl")


def test_boidsim_grid_fp64_mode_tracks_fp64_oracle(tmp_path):
    # BOIDS_ENGINE_MODE=grid_f64 runs the FP64 grid path behind the same
    # executable: it follows the FP64 oracle from the grid engine's init
    out = run("boidsim_grid", "-s", TESTINI, "-T", "20", "-N", "4000", "-F", "1.2", "-o", "g.ptcl",
              "--velocity-output", "gv.ptcl", cwd=tmp_path, env={"BOIDS_ENGINE_MODE": "grid_f64"})
    assert "grid num:" in out
    fr = ptcl.PtclFile(tmp_path / "g.ptcl").frames()
    P = O.params(1.2, *TESTINI_RULES)
    p32, v32 = O.init_reference("test", 4000, 1.2, 0.003, 0.001, seed=12345)   # grid engine seeds srand48
    assert np.array_equal(fr[0], p32)
    p, v = p32.astype(np.float64), v32.astype(np.float64)
    for t in range(1, 20):
        O.step(P, p, v, enumeration=O.ENUM_CELLS)
        d = fr[t].astype(np.float64) - p
        d -= 1.2 * np.round(d / 1.2)
        assert np.abs(d).max() < 1e-6, t
