"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs the reference engines compiled by oracle/Makefile.ref (oracle/_ref/
ref_allpairs = boidsim's src/boid_simulation.cpp, oracle/_ref/ref_grid =
boidsim_grid's src/boid_simulation_grid.cpp) through their public
BoidSimulation API and stores inputs and expected outputs as .npz data.
Needs /root/reference (this container only); the fixtures it writes are
committed and travel without the reference.

    python tests/golden/make_golden.py               # every fixture
    python tests/golden/make_golden.py grid_testini  # only ref_grid_testini_n1000.npz
    python tests/golden/make_golden.py grid_cfg1     # only ref_grid_cfg1_n16384.npz
"""
from __future__ import annotations

import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

# cfg0 rules of the workloads: radii 0.02/0.05/0.08, view 270 deg (a = 0.75),
# weights 0.002/0.06/0.008, |v| in [0.001, 0.005]
CFG0 = dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001)
# settings/test.ini (reference): N=1000 T=30 INIT=test RANDOM_SEED=12345
TESTINI = dict(sep=(0.1, 0.9, 0.0001), ali=(0.1, 0.7, 0.1), coh=(0.1, 0.8, 0.1), vmax=0.003, vmin=0.001)
# settings/example_grid.ini (reference): distance-only rules (a = 1)
EXGRID = dict(sep=(0.05, 1.0, 0.0001), ali=(0.05, 1.0, 0.1), coh=(0.05, 1.0, 0.1), vmax=0.003, vmin=0.001)


def run_ref(binary, n, T, F, rules, init, seed, dump):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "dump.bin")
        args = [str(x) for x in (n, T, F, *rules["sep"], *rules["ali"], *rules["coh"], rules["vmax"], rules["vmin"])]
        subprocess.run([os.path.join(O.REF_DIR, binary), out, *args, init, str(seed),
                        ",".join(str(t) for t in dump)], check=True, stdout=subprocess.DEVNULL)
        return O.read_ref_dump(out, n)


def ptcl_bytes(frames, n, T, F):
    """.ptcl v0.2 as boidsim writes it: 50-byte header (src/dtype.h:19-36,
    src/main.cpp:60-74) then positions of states 0..T-1 (src/main.cpp:187-292)."""
    hdr = struct.pack("<4sBBIIIIIffffff", b"PTCL", 0, 2, 50, n, T, 0, 30,
                      0.0, F, 0.0, F, 0.0, F)
    return hdr + b"".join(frames[t].astype("<f4").tobytes() for t in range(T))


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def rules_array(F, r):
    return np.array([F, *r["sep"], *r["ali"], *r["coh"], r["vmax"], r["vmin"]], np.float64)


def grid_testini():
    """6. boidsim_grid with settings/test.ini: the whole .ptcl file (positions
    and velocities) as a digest.  The grid engine seeds srand48(12345) and
    sizes GRID_NUM = 4 here (src/boid_simulation_grid.cpp:80-89)."""
    n, T, F = 1000, 30, 1.0
    d = run_ref("ref_grid", n, T, F, TESTINI, "test", 12345, list(range(T + 1)))
    pos = {t: d[t][0] for t in d}
    vel = {t: d[t][1] for t in d}
    save("ref_grid_testini_n1000.npz", rules=rules_array(F, TESTINI),
         pos0=d[0][0], vel0=d[0][1], pos30=d[30][0], vel30=d[30][1],
         ptcl_sha256=np.frombuffer(hashlib.sha256(ptcl_bytes(pos, n, T, F)).digest(), np.uint8),
         vptcl_sha256=np.frombuffer(hashlib.sha256(ptcl_bytes(vel, n, T, F)).digest(), np.uint8))


def grid_cfg1():
    """7. boidsim_grid at full cfg1 size: N = 16384, F = 1, cfg0 rules, the
    reference's random_uniform init after srand48(1), 200 updates.  Stored as
    digests only (the fixture stays small): SHA-256 of the initial state and
    of the .ptcl files of positions and velocities (states 0..199)."""
    n, T, F = 16384, 200, 1.0
    d = run_ref("ref_grid", n, T, F, CFG0, "random_uniform", 1, list(range(T + 1)))
    pos = {t: d[t][0] for t in d}
    vel = {t: d[t][1] for t in d}
    init = np.concatenate([d[0][0], d[0][1]], 1).astype("<f4")
    save("ref_grid_cfg1_n16384.npz", rules=rules_array(F, CFG0), steps=np.int64(T),
         init_sha256=np.frombuffer(hashlib.sha256(init.tobytes()).digest(), np.uint8),
         ptcl_sha256=np.frombuffer(hashlib.sha256(ptcl_bytes(pos, n, T, F)).digest(), np.uint8),
         vptcl_sha256=np.frombuffer(hashlib.sha256(ptcl_bytes(vel, n, T, F)).digest(), np.uint8),
         final_sha256=np.frombuffer(hashlib.sha256(
             np.concatenate([d[T][0], d[T][1]], 1).astype("<f4").tobytes()).digest(), np.uint8))


def main():
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile.ref")], check=True)
    if sys.argv[1:] == ["grid_testini"]:      # one fixture only (the others stay byte-identical)
        grid_testini()
        return
    if sys.argv[1:] == ["grid_cfg1"]:
        grid_cfg1()
        return

    # 1. boidsim, random_uniform init (drand48 never seeded), cfg0 rules
    n, T, F = 1024, 100, 1.0
    d = run_ref("ref_allpairs", n, T, F, CFG0, "random_uniform", 0, [0, 1, 100])
    save("ref_allpairs_ru_n1024.npz", rules=rules_array(F, CFG0),
         pos0=d[0][0], vel0=d[0][1], pos1=d[1][0], vel1=d[1][1], force1=d[1][2],
         pos100=d[100][0], vel100=d[100][1], force100=d[100][2])

    # 2. boidsim_grid, random_uniform init with srand48(1), cfg0 rules, GRID_NUM 6
    n, T, F = 2048, 50, 1.0
    d = run_ref("ref_grid", n, T, F, CFG0, "random_uniform", 1, [0, 50])
    save("ref_grid_ru_seed1_n2048.npz", rules=rules_array(F, CFG0),
         pos0=d[0][0], vel0=d[0][1], pos50=d[50][0], vel50=d[50][1])

    # 3. boidsim with settings/test.ini: the whole .ptcl file as a digest
    n, T, F = 1000, 30, 1.0
    d = run_ref("ref_allpairs", n, T, F, TESTINI, "test", 12345, list(range(T + 1)))
    pos = {t: d[t][0] for t in d}
    vel = {t: d[t][1] for t in d}
    save("ref_allpairs_testini_n1000.npz", rules=rules_array(F, TESTINI),
         pos0=d[0][0], vel0=d[0][1], pos30=d[30][0], vel30=d[30][1],
         ptcl_sha256=np.frombuffer(hashlib.sha256(ptcl_bytes(pos, n, T, F)).digest(), np.uint8),
         vptcl_sha256=np.frombuffer(hashlib.sha256(ptcl_bytes(vel, n, T, F)).digest(), np.uint8))

    # 4. boidsim with a file init (N x 6 float32): generic isotropic state
    n, T, F = 1024, 100, 1.0
    rng = np.random.default_rng(20240611)
    x = rng.random((n, 3)) * F
    sp = 0.001 + 0.004 * rng.random(n)
    ct = 2 * rng.random(n) - 1
    ph = 2 * np.pi * rng.random(n)
    st = np.sqrt(1 - ct * ct)
    v = sp[:, None] * np.stack([st * np.cos(ph), st * np.sin(ph), ct], 1)
    init = np.concatenate([x, v], 1).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        fpath = os.path.join(td, "init.bin")
        init.tofile(fpath)
        d = run_ref("ref_allpairs", n, T, F, CFG0, fpath, 0, [0, 1, 10, 100])
        dg = run_ref("ref_grid", n, T, F, CFG0, fpath, 0, [100])
    assert np.array_equal(d[100][0], dg[100][0]) and np.array_equal(d[100][1], dg[100][1]), \
        "reference grid and all-pairs disagree on a shared init"
    save("ref_allpairs_file_n1024.npz", rules=rules_array(F, CFG0), init=init,
         pos1=d[1][0], vel1=d[1][1], force1=d[1][2], pos10=d[10][0], vel10=d[10][1],
         pos100=d[100][0], vel100=d[100][1], force100=d[100][2])

    # 5. boidsim_grid with settings/example_grid.ini rules (distance only)
    n, T, F = 2048, 20, 1.0
    d = run_ref("ref_grid", n, T, F, EXGRID, "random_uniform", 12345, [0, 20])
    save("ref_grid_distonly_n2048.npz", rules=rules_array(F, EXGRID),
         pos0=d[0][0], vel0=d[0][1], pos20=d[20][0], vel20=d[20][1])

    grid_testini()
    grid_cfg1()


if __name__ == "__main__":
    main()
