"""CPU tests of the drop-in host layer: the Boost-free .ini/CLI parser
(restating src/args.cpp:17-153) and the .ptcl tooling (restating
tools/*.py), including the split/resume invariant of test/compare_split.sh
with the oracle as the engine."""
import os
import struct
import subprocess

import numpy as np
import pytest

from boidssimulatormpi_amd import ptcl
from oracle import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CSRC = os.path.join(ROOT, "boidssimulatormpi_amd", "csrc")
TESTINI = os.path.join(ROOT, "tests", "golden", "settings_test.ini")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "args_probe")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, "-o", exe,
                    os.path.join(ROOT, "tests", "args_probe.cpp"), os.path.join(CSRC, "args.cpp")], check=True)
    return exe


def parse(probe, *args):
    r = subprocess.run([probe, *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(line.split("=", 1) for line in r.stdout.strip().splitlines())


def test_ini_values(probe):
    a = parse(probe, "-s", TESTINI, "-o", "out.ptcl")
    assert a["T"] == "30" and a["N"] == "1000" and float(a["F"]) == 1.0
    assert a["init"] == "test" and a["seed"] == "12345" and a["output"] == "out.ptcl"
    assert [float(x) for x in a["sep"].split(",")] == [0.1, 0.9, 0.0001]
    assert [float(x) for x in a["ali"].split(",")] == [0.1, 0.7, 0.1]
    assert [float(x) for x in a["coh"].split(",")] == [0.1, 0.8, 0.1]
    assert float(a["vmax"]) == 0.003 and float(a["vmin"]) == 0.001


def test_cli_overrides_ini_and_forms(probe):
    a = parse(probe, "--setting=" + TESTINI, "-N2048", "-T", "7", "--field-size", "0.5", "--vmin=0.002",
              "--vmax", "0.004", "--random-seed", "-3", "-i", "random_uniform", "--velocity-output", "v.ptcl",
              "--force-data-output")
    assert a["N"] == "2048" and a["T"] == "7" and float(a["F"]) == 0.5
    assert float(a["vmin"]) == 0.002 and float(a["vmax"]) == 0.004 and a["seed"] == "-3"
    assert a["init"] == "random_uniform" and a["velocity_output"] == "v.ptcl" and a["force"] == "1"
    assert a["output"] == ""          # -i no longer aliases the output file name
    assert [float(x) for x in a["coh"].split(",")] == [0.1, 0.8, 0.1]   # untouched ini values


def test_ini_comments_and_bad_values(probe, tmp_path):
    ini = tmp_path / "x.ini"
    ini.write_text("# comment\n; other comment\n[Global]\nPOPULATION = 12x\nFIELD_SIZE=  2.5  \n"
                   "[Cohesion]\nSIGHT_ANGLE= 0.3333333333333333\n")
    a = parse(probe, "-s", str(ini))
    assert a["N"] == "0"              # not a complete unsigned: ignored, as get_optional<unsigned> does
    assert float(a["F"]) == 2.5
    assert float(a["coh"].split(",")[1]) == 0.3333333333333333


def test_bad_options_exit(probe):
    assert subprocess.run([probe, "--nope"], capture_output=True).returncode == 255
    assert subprocess.run([probe, "-N", "abc"], capture_output=True).returncode == 255
    assert subprocess.run([probe, "-N"], capture_output=True).returncode == 255
    r = subprocess.run([probe, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "--setting" in r.stdout


# ------------------------------------------------------------------ .ptcl ---
def test_header_layout(tmp_path):
    f = tmp_path / "a.ptcl"
    fr = np.random.default_rng(0).random((4, 10, 3)).astype(np.float32)
    ptcl.write(f, fr, 2.0, t_0=3)
    raw = f.read_bytes()
    assert len(raw) == 50 + 4 * 10 * 12
    hdr = struct.unpack("<4sBBIIIIIffffff", raw[:50])
    assert hdr == (b"PTCL", 0, 2, 50, 10, 4, 3, 30, 0.0, 2.0, 0.0, 2.0, 0.0, 2.0)
    p = ptcl.PtclFile(f)
    assert p.check_metadata() and np.array_equal(p.frames(), fr)


def test_fix_split_merge(tmp_path):
    fr = np.random.default_rng(1).random((10, 7, 3)).astype(np.float32)
    a = tmp_path / "a.ptcl"
    ptcl.write(a, fr, 1.0)
    with open(a, "ab") as f:                       # a truncated, half-written frame
        f.write(b"\x00" * 40)
    with pytest.raises(ValueError):
        ptcl.PtclFile(a).check_metadata()
    ptcl.fix(a)
    assert ptcl.PtclFile(a).check_metadata() and np.array_equal(ptcl.PtclFile(a).frames(), fr)
    b, c = tmp_path / "b.ptcl", tmp_path / "c.ptcl"
    ptcl.split(a, b, 0, 4)
    ptcl.split(a, c, 4, 6)
    assert ptcl.PtclFile(c).header.t_0 == 4
    m = tmp_path / "m.ptcl"
    ptcl.merge([b, c], m)
    assert m.read_bytes() == a.read_bytes()


def test_split_resume_equals_continuous(tmp_path):
    # test/compare_split.sh with the oracle as the engine: 30 steps straight
    # vs 15 steps, init file from frame 14 (pos + vel), 16 more steps, split, merge
    P = O.params(1.0, (0.1, 0.9, 0.0001), (0.1, 0.7, 0.1), (0.1, 0.8, 0.1), 0.003, 0.001)
    p0, v0 = O.init_reference("test", 300, 1.0, 0.003, 0.001)

    def run(p, v, T):
        fp, fv = [], []
        for _ in range(T):
            fp.append(p.copy()); fv.append(v.copy())
            O.step(P, p, v)
        return np.array(fp), np.array(fv)

    full = tmp_path / "full.ptcl"
    ptcl.write(full, run(p0.copy(), v0.copy(), 30)[0], 1.0)
    a, av = tmp_path / "a.ptcl", tmp_path / "av.ptcl"
    fp, fv = run(p0.copy(), v0.copy(), 15)
    ptcl.write(a, fp, 1.0)
    ptcl.write(av, fv, 1.0)
    init = tmp_path / "init.bin"
    ptcl.generate_initialization_file(a, av, 14, init)
    st = np.fromfile(init, "<f4").reshape(-1, 6)
    b = tmp_path / "b.ptcl"
    ptcl.write(b, run(st[:, :3].copy(), st[:, 3:].copy(), 16)[0], 1.0)
    tmp = tmp_path / "tmp.ptcl"
    ptcl.split(b, tmp, 1, 15)
    merged = tmp_path / "merged.ptcl"
    ptcl.merge([a, tmp], merged)
    assert merged.read_bytes() == full.read_bytes()


def test_calc_vel_unwraps_periodic_jumps(tmp_path):
    fr = np.array([[[0.99, 0.5, 0.5]], [[0.002, 0.5, 0.5]], [[0.014, 0.5, 0.5]]], np.float32)
    f = tmp_path / "v.ptcl"
    ptcl.write(f, fr, 1.0)
    x0, v = ptcl.calc_vel(ptcl.PtclFile(f), 0, step=2)
    assert np.allclose(v[0], [(1.014 - 0.99) / 2, 0, 0], atol=1e-6)
