"""CPU tests of the sweep front-end (boidssimulatormpi_amd/sweep.py): its
.ini reader agrees with the drop-in driver's C++ parser, the job list
follows tools/gen_jobscript.py:20-21,92-111, the synthetic many-flock sweep
has the stated distribution, and bs_create_batch rejects bad flock tables
before touching a device."""
import os
import subprocess

import numpy as np
import pytest

from boidssimulatormpi_amd import engine as E
from boidssimulatormpi_amd import sweep as S

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "boidssimulatormpi_amd", "csrc")
TESTINI = os.path.join(ROOT, "tests", "golden", "settings_test.ini")

VEL_INI = """[Global]
# DENSITY is 16384 (2**14)
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
MIN = 0.0010
MAX = 0.0030
"""

ODD_INI = "# c\n; c2\n[Global]\nPOPULATION = 12x\nFIELD_SIZE=  2.5  \nTIME_STEP = 7\nRANDOM_SEED = -4\n" \
          "[Velocity]\nMAX = 1e-3\nMIN = .5e-3\n[Cohesion]\nSIGHT_ANGLE= 0.3333333333333333\n" \
          "SIGHT_DISTANCE = 0.1.2\n"


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "args_probe")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, "-o", exe,
                    os.path.join(ROOT, "tests", "args_probe.cpp"), os.path.join(CSRC, "args.cpp")], check=True)
    return exe


def cpp_settings(probe, path):
    r = subprocess.run([probe, "-s", path], capture_output=True, text=True)
    assert r.returncode == 0
    a = dict(line.split("=", 1) for line in r.stdout.strip().splitlines())
    t = lambda k: tuple(float(x) for x in a[k].split(","))
    return {"time_step": int(a["T"]), "population": int(a["N"]), "field_size": float(a["F"]),
            "init": a["init"], "seed": int(a["seed"]), "sep": t("sep"), "ali": t("ali"), "coh": t("coh"),
            "vmax": float(a["vmax"]), "vmin": float(a["vmin"])}


@pytest.mark.parametrize("text", [None, VEL_INI, ODD_INI])
def test_ini_reader_matches_driver(probe, tmp_path, text):
    path = TESTINI
    if text is not None:
        path = str(tmp_path / "s.ini")
        open(path, "w").write(text)
    assert S.read_settings(path) == cpp_settings(probe, path)


def test_populations_and_jobs(tmp_path):
    assert S.populations(["2**10", "[2048, 4096]", "[range(1, 3), 5]"]) == [1024, 2048, 4096, 1, 2, 5]
    with pytest.raises(ValueError):
        S.populations(["__import__('os')"])
    ini = tmp_path / "vmin0.0010_vmax0.0030.ini"
    ini.write_text(VEL_INI)
    jobs = S.jobs_from_settings([str(ini)], [2048, 16384], 16384.0, 100)
    assert [j.name for j in jobs] == ["vmin0.0010_vmax0.0030_N0002048_T100", "vmin0.0010_vmax0.0030_N0016384_T100"]
    assert jobs[0].field == (2048 / 16384.0) ** (1 / 3) and jobs[1].field == 1.0
    assert jobs[0].sep == (0.01, 0.5, 0.002) and jobs[0].vmax == 0.003 and jobs[0].vmin == 0.001
    assert jobs[0].init == "random_uniform" and jobs[0].seed == 12345


def test_generated_sweep_distribution():
    jobs = S.generated_sweep(512, 3)
    ns = np.array([j.n for j in jobs])
    assert len(jobs) == 512 and ns.min() >= 1024 and ns.max() < 65536
    # log2 N ~ U[10, 16]: mean 13, sd 6/sqrt(12)
    l2 = np.log2(ns)
    assert abs(l2.mean() - 13.0) < 0.3 and abs(l2.std() - 6 / 12 ** 0.5) < 0.2
    for j in jobs:
        assert abs(j.n / j.field ** 3 - 16384.0) < 1e-6 * j.n
        assert 0.002 <= j.vmax <= 0.01 and j.vmin == j.vmax / 5
        s = j.coh[0] / 0.08
        assert 0.7 <= s <= 1.3 and j.sep[0] == pytest.approx(0.02 * s) and j.ali[0] == pytest.approx(0.05 * s)
    assert [j.n for j in S.generated_sweep(512, 3)] == list(ns)      # deterministic
    # every flock has at least 3 cells per axis at its largest radius
    for j in jobs:
        assert np.floor(j.field / (j.coh[0] * 1.0001)) >= 3


def _batch_error(params):
    arr = (E.Params * len(params))(*params)
    import ctypes
    h = ctypes.c_void_p()
    rc = E.lib().bs_create_batch(arr, len(params), ctypes.byref(h))
    return rc, E.lib().bs_last_error(None).decode()


def test_create_batch_rejects_bad_tables():
    ok = E.make_params(1000, 1.0, mode=E.MODE_GRID_F64)
    rc, msg = _batch_error([E.make_params(1000, 1.0, mode=E.MODE_ALLPAIRS_F64)])
    assert rc == -1 and "grid modes" in msg
    rc, msg = _batch_error([ok, E.make_params(1000, 0.2, mode=E.MODE_GRID_F64)])
    assert rc == -1 and "flock 1" in msg and "3 cells" in msg
    rc, msg = _batch_error([ok, E.make_params(1000, 1.0, mode=E.MODE_GRID_MIXED)])
    assert rc == -1 and "same" in msg
    rc, msg = _batch_error([ok, E.make_params(1000, 1.0, mode=E.MODE_GRID_F64, flags=E.FLAG_DIGEST)])
    assert rc == -1 and "same" in msg
    rc, msg = _batch_error([ok, E.make_params(1000, float("nan"), mode=E.MODE_GRID_F64)])
    assert rc == -1 and "flock 1" in msg
    import ctypes
    h = ctypes.c_void_p()
    assert E.lib().bs_create_batch(None, 0, ctypes.byref(h)) == -1


def test_autotune_candidates_and_cache_key():
    from boidssimulatormpi_amd import autotune as T
    p = E.make_params(1 << 20, 4.0, mode=E.MODE_GRID_MIXED)
    assert T.auto_cells(p) == 49          # floor(4 / (0.08 * 1.0001))
    c = list(T.candidates(p))
    assert len(c) == len(T.BLOCKS) * len(T.TILES) * len(T.FILLS) * len(T.CELL_SCALES)
    for t, ps in c:
        q = ps[0]
        assert (q.block, q.tile, q.fill) == (t["block"], t["tile"], t["fill"])
        assert q.cells_per_axis == (0 if t["cell_scale"] == 1.0 else int(49 * t["cell_scale"]))
        assert q.n == p.n and q.field == p.field and q.mode == p.mode
    # a cell scale that would leave < 3 cells per axis is dropped
    small = E.make_params(1000, 0.3, mode=E.MODE_GRID_F64)
    assert T.auto_cells(small) == 3
    # the cells-per-boid cap is the exact integer cube root (perfect cubes included)
    for n, g in ((64, 4), (4096, 16), (32768, 32), (1 << 18, 64), (4095, 15), (35937, 33), (35936, 32)):
        assert T.icbrt(n) == g
        assert T.auto_cells(E.make_params(n, 20.0, mode=E.MODE_GRID_F64)) == g
    assert all(t["cell_scale"] == 1.0 for t, _ in T.candidates(small))
    assert T.workload_key(p) == T.workload_key(E.make_params(1 << 20, 4.0, mode=E.MODE_GRID_MIXED))
    assert T.workload_key(p) != T.workload_key(E.make_params(1 << 20, 4.1, mode=E.MODE_GRID_MIXED))
    assert T.workload_key([p, small]).startswith("mode3_flocks2_")
