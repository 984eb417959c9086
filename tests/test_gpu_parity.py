"""Parity of the HIP engine (through its C ABI) with the reference and the oracle.

Bars:
  * float mode (BS_MODE_ALLPAIRS_F32REF): bit-identical to the reference's
    own outputs (golden fixtures written by the compiled reference).
  * FP64 modes: neighbour sets bit-exact with the FP64 oracle every step
    (digests, GPU state fed to the oracle), positions/velocities within
    1e-12 after one step and 1e-9 after ten (errors scaled by F and vmax).
  * mixed mode: neighbour sets bit-exact with the FP64 oracle every step;
    per-step error against FP64 within 1e-6 (scaled by F and vmax).
"""
import hashlib
import os
import struct
import sys

import numpy as np
import pytest

import boidssimulatormpi_amd as B
from boidssimulatormpi_amd import engine as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG0 = dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001)
SWEEP = dict(sep=(0.01, 0.5, 0.002), ali=(0.05, 1 / 3, 0.06), coh=(0.05, 0.5, 0.008), vmax=0.005, vmin=0.001)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rules_of(arr):
    F, sr, sa, sw, ar, aa, aw, cr, ca, cw, vmax, vmin = [float(x) for x in arr]
    return F, dict(sep=(sr, sa, sw), ali=(ar, aa, aw), coh=(cr, ca, cw), vmax=vmax, vmin=vmin)


def make(n, F, rules, mode, flags=0, cells=0):
    say(f"engine n={n} F={F} mode={E.MODE_NAMES[mode]} flags={flags}")
    return B.Engine(B.make_params(n, F, **rules, mode=mode, flags=flags, cells_per_axis=cells))


def wvec(rules):
    return np.float32([rules["coh"][2]] * 3 + [rules["sep"][2]] * 3 + [rules["ali"][2]] * 3)


def scaled_err(pa, va, pb, vb, F, vmax):
    dp = pa - pb
    dp -= F * np.round(dp / F)          # a wrap decided differently is the same point
    return float(np.abs(dp).max() / F), float(np.abs(va - vb).max() / vmax)


def iso_state(n, F, seed, rules=CFG0):
    return B.generate_init(E.INIT_ISOTROPIC, n, F, rules["vmax"], rules["vmin"], seed=seed)


# ------------------------------------------------ float mode vs reference ---
def test_f32ref_random_uniform(golden_dir):
    g = np.load(os.path.join(golden_dir, "ref_allpairs_ru_n1024.npz"))
    F, rules = rules_of(g["rules"])
    e = make(1024, F, rules, E.MODE_ALLPAIRS_F32REF, E.FLAG_FORCES)
    e.set_state(xyzv6=np.concatenate([g["pos0"], g["vel0"]], 1))
    for t in range(1, 101):
        e.step(1)
        if t in (1, 100):
            p, v = e.state_f32()
            assert np.array_equal(p, g[f"pos{t}"]) and np.array_equal(v, g[f"vel{t}"]), t
            assert np.array_equal(e.forces_f32() * wvec(rules), g[f"force{t}"]), t


def test_f32ref_file_init(golden_dir):
    g = np.load(os.path.join(golden_dir, "ref_allpairs_file_n1024.npz"))
    F, rules = rules_of(g["rules"])
    e = make(1024, F, rules, E.MODE_ALLPAIRS_F32REF, E.FLAG_FORCES)
    e.set_state(xyzv6=g["init"])
    for t in range(1, 101):
        e.step(1)
        if t in (1, 10, 100):
            p, v = e.state_f32()
            assert np.array_equal(p, g[f"pos{t}"]) and np.array_equal(v, g[f"vel{t}"]), t
        if t in (1, 100):
            assert np.array_equal(e.forces_f32() * wvec(rules), g[f"force{t}"]), t


@pytest.mark.parametrize("name,steps", [("ref_grid_ru_seed1_n2048.npz", 50), ("ref_grid_distonly_n2048.npz", 20)])
def test_f32ref_matches_reference_grid(golden_dir, name, steps):
    g = np.load(os.path.join(golden_dir, name))
    F, rules = rules_of(g["rules"])
    e = make(2048, F, rules, E.MODE_ALLPAIRS_F32REF)
    e.set_state(xyzv6=np.concatenate([g["pos0"], g["vel0"]], 1))
    e.step(steps)
    p, v = e.state_f32()
    assert np.array_equal(p, g[f"pos{steps}"]) and np.array_equal(v, g[f"vel{steps}"])


def test_f32ref_testini_ptcl(golden_dir):
    g = np.load(os.path.join(golden_dir, "ref_allpairs_testini_n1000.npz"))
    F, rules = rules_of(g["rules"])
    e = make(1000, F, rules, E.MODE_ALLPAIRS_F32REF)
    e.set_state(xyzv6=np.concatenate([g["pos0"], g["vel0"]], 1))
    hdr = struct.pack("<4sBBIIIIIffffff", b"PTCL", 0, 2, 50, 1000, 30, 0, 30, 0.0, F, 0.0, F, 0.0, F)
    hp, hv = hashlib.sha256(hdr), hashlib.sha256(hdr)
    for _ in range(30):
        p, v = e.state_f32()
        hp.update(p.astype("<f4").tobytes()); hv.update(v.astype("<f4").tobytes())
        e.step(1)
    assert hp.digest() == g["ptcl_sha256"].tobytes() and hv.digest() == g["vptcl_sha256"].tobytes()


# ------------------------------------------ float grid mode vs reference ---
@pytest.mark.parametrize("name,steps", [("ref_grid_ru_seed1_n2048.npz", 50), ("ref_grid_distonly_n2048.npz", 20),
                                        ("ref_grid_testini_n1000.npz", 30)])
def test_f32_grid_matches_reference_grid(golden_dir, name, steps):
    # BS_MODE_GRID_F32REF is boidsim_grid's arithmetic: bit-identical to the
    # reference's own grid engine on its own fixtures
    g = np.load(os.path.join(golden_dir, name))
    F, rules = rules_of(g["rules"])
    n = g["pos0"].shape[0]
    e = make(n, F, rules, E.MODE_GRID_F32REF)
    assert e.info().mode == E.MODE_GRID_F32REF and e.info().cells_per_axis >= 3
    e.set_state(xyzv6=np.concatenate([g["pos0"], g["vel0"]], 1))
    e.step(steps)
    p, v = e.state_f32()
    assert np.array_equal(p, g[f"pos{steps}"]) and np.array_equal(v, g[f"vel{steps}"])


@pytest.mark.parametrize("mode", [E.MODE_GRID_F32REF, E.MODE_ALLPAIRS_F32REF])
def test_f32_cfg1_full_size_matches_reference_ptcl(golden_dir, mode):
    # boidsim_grid at the full cfg1 size (N = 16384, F = 1, its random_uniform
    # init after srand48(1)): the reference grid engine's .ptcl digests of
    # positions and velocities over 200 updates, byte for byte (boidsim's
    # all-pairs arithmetic gives the same bytes)
    g = np.load(os.path.join(golden_dir, "ref_grid_cfg1_n16384.npz"))
    F, rules = rules_of(g["rules"])
    n, T = 16384, int(g["steps"])
    st = B.generate_init(E.INIT_RANDOM_UNIFORM, n, F, rules["vmax"], rules["vmin"], seed=1)
    assert hashlib.sha256(st.astype("<f4").tobytes()).digest() == g["init_sha256"].tobytes()
    e = make(n, F, rules, mode)
    assert e.info().mode == mode
    e.set_state(xyzv6=st)
    hdr = struct.pack("<4sBBIIIIIffffff", b"PTCL", 0, 2, 50, n, T, 0, 30, 0.0, F, 0.0, F, 0.0, F)
    hp, hv = hashlib.sha256(hdr), hashlib.sha256(hdr)
    for _ in range(T):
        p, v = e.state_f32()
        hp.update(p.astype("<f4").tobytes()); hv.update(v.astype("<f4").tobytes())
        e.step(1)
    p, v = e.state_f32()
    assert hp.digest() == g["ptcl_sha256"].tobytes() and hv.digest() == g["vptcl_sha256"].tobytes()
    assert hashlib.sha256(np.concatenate([p, v], 1).astype("<f4").tobytes()).digest() == g["final_sha256"].tobytes()


def test_f32_grid_file_init_forces(golden_dir):
    g = np.load(os.path.join(golden_dir, "ref_allpairs_file_n1024.npz"))
    F, rules = rules_of(g["rules"])
    e = make(1024, F, rules, E.MODE_GRID_F32REF, E.FLAG_FORCES)
    assert e.info().mode == E.MODE_GRID_F32REF
    e.set_state(xyzv6=g["init"])
    for t in range(1, 101):
        e.step(1)
        if t in (1, 10, 100):
            p, v = e.state_f32()
            assert np.array_equal(p, g[f"pos{t}"]) and np.array_equal(v, g[f"vel{t}"]), t
        if t in (1, 100):
            assert np.array_equal(e.forces_f32() * wvec(rules), g[f"force{t}"]), t


@pytest.mark.parametrize("n,F,cells", [(16384, 1.0, 0), (3000, 0.25, 3), (20000, 1.6, 0), (16384, 0.62, 0)])
def test_f32_grid_equals_allpairs(n, F, cells):
    # the float grid equals the float all-pairs step bit for bit over a long
    # run (chaotic: any difference in a sum's order would grow), with its
    # neighbour digests and forces; also on the smallest grid (3 cells per
    # axis).  Boids within the largest radius per owner: ~35 at N = 16384,
    # F = 1 (sorted in one register), ~150 at F = 0.62 (two and four
    # registers), ~400 at N = 3000, F = 0.25 (sorted in LDS)
    st = iso_state(n, F, 3)
    a = make(n, F, CFG0, E.MODE_ALLPAIRS_F32REF, E.FLAG_DIGEST | E.FLAG_FORCES)
    b = make(n, F, CFG0, E.MODE_GRID_F32REF, E.FLAG_DIGEST | E.FLAG_FORCES, cells=cells)
    assert b.info().mode == E.MODE_GRID_F32REF and (cells == 0 or b.info().cells_per_axis == cells)
    a.set_state(xyzv6=st)
    b.set_state(xyzv6=st)
    for t in range(4):
        steps = 1 if t < 3 else 150
        a.step(steps)
        b.step(steps)
        pa, va = a.state_f32()
        pb, vb = b.state_f32()
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), t
        assert np.array_equal(a.digest(), b.digest()), t
        assert np.array_equal(a.forces_f32(), b.forces_f32()), t
    say(f"float grid == float all-pairs over 153 updates, N={n}, {b.info().cells_per_axis} cells per axis")


def test_f32_grid_crowded_owners_and_oracle():
    # Two crowded cubes of edge 0.01 (every pair inside a cube lies within
    # every radius).  Cube A holds 1500 boids: each of their owners has 1499
    # boids within the largest radius, more than the per-wave sorted list
    # holds (1024), and takes the serial 27-way merge.  Cube B holds 900:
    # their owners sort lists of ~900 ids (1024 with the padding).  Everyone
    # else sorts short lists.  All must equal the float all-pairs step and the
    # float oracle.
    n, F = 8000, 1.0
    rng = np.random.default_rng(21)
    st = iso_state(n, F, 5)
    perm = rng.permutation(n)
    sel_a, sel_b = perm[:1500], perm[1500:2400]
    st[sel_a, 0:3] = (0.4 + 0.01 * rng.random((1500, 3))).astype(np.float32)
    st[sel_b, 0:3] = (0.7 + 0.01 * rng.random((900, 3))).astype(np.float32)
    a = make(n, F, CFG0, E.MODE_ALLPAIRS_F32REF, E.FLAG_DIGEST | E.FLAG_FORCES)
    b = make(n, F, CFG0, E.MODE_GRID_F32REF, E.FLAG_DIGEST | E.FLAG_FORCES)
    assert b.info().mode == E.MODE_GRID_F32REF
    a.set_state(xyzv6=st)
    b.set_state(xyzv6=st)
    P = O.params(F, **CFG0)
    p, v = st[:, :3].copy(), st[:, 3:].copy()
    _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
    # the cubes are as crowded as intended: cohesion counts (view cone
    # included) above 1024 in A and above 512 in B
    assert int(((d[sel_a, 0] & 0x1FFFFF) > 1024).sum()) > 500
    assert int(((d[sel_b, 0] & 0x1FFFFF) > 512).sum()) > 300
    for t in range(3):
        a.step(1)
        b.step(1)
        pa, va = a.state_f32()
        pb, vb = b.state_f32()
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), t
        assert np.array_equal(a.digest(), b.digest()) and np.array_equal(a.forces_f32(), b.forces_f32()), t
        if t == 0:
            assert np.array_equal(pb, p) and np.array_equal(vb, v) and np.array_equal(b.digest(), d)


# ------------------------------------------------------- FP64 all-pairs ---
def test_f64_allpairs_bit_exact_with_oracle():
    n, F = 1024, 1.0
    st = iso_state(n, F, 0)
    e = make(n, F, CFG0, E.MODE_ALLPAIRS_F64, E.FLAG_FORCES | E.FLAG_DIGEST)
    e.set_state(xyzv6=st)
    P = O.params(F, **CFG0)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    for t in range(10):
        f, d = O.step(P, p, v, forces=True, digest=True)
        e.step(1)
        gp, gv = e.state_f64()
        assert np.array_equal(e.digest(), d), t
        assert np.array_equal(gp, p) and np.array_equal(gv, v), t   # same order of sums: same bits
        assert np.array_equal(e.forces_f64(), f), t


# -------------------------------------------------------------- grid ---
@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
@pytest.mark.parametrize("n,F,rules", [(16384, 1.0, CFG0), (4096, 1.0, SWEEP), (20000, 1.6, CFG0)])
def test_grid_neighbour_sets_each_step(mode, n, F, rules):
    st = iso_state(n, F, 1, rules)
    e = make(n, F, rules, mode, E.FLAG_DIGEST | E.FLAG_FORCES)
    assert e.info().mode == mode and e.info().cells_per_axis >= 3
    e.set_state(xyzv6=st)
    P = O.params(F, **rules)
    gp, gv = e.state_f64()
    worst = (0.0, 0.0)
    for t in range(6):
        p, v = gp.copy(), gv.copy()
        f, d = O.step(P, p, v, forces=True, digest=True, enumeration=O.ENUM_CELLS)
        e.step(1)
        gd = e.digest()
        bad = np.nonzero((gd != d).any(1))[0]
        assert bad.size == 0, f"step {t}: {bad.size} boids with different neighbour sets, e.g. {bad[:5]}"
        gp, gv = e.state_f64()
        ep, ev = scaled_err(gp, gv, p, v, F, rules["vmax"])
        worst = (max(worst[0], ep), max(worst[1], ev))
        tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
        assert ep <= tol and ev <= tol, (t, ep, ev)
    say(f"per-step max scaled error pos={worst[0]:.3e} vel={worst[1]:.3e} rechecks={e.rechecks()}")


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
@pytest.mark.parametrize("F,cells", [(0.25, 3), (0.33, 4)])
def test_smallest_grids_wrapped_tiles(mode, F, cells):
    # 3 or 4 cells per axis, x binned at half the cell edge: 6 or 8 x columns,
    # so a chunk's tile (its columns plus two on either side) wraps around the
    # row and can hold a column twice, while each boid's 5-column window
    # (distinct columns) must still see every neighbour exactly once.  Dense:
    # ~400 cohesion neighbours per boid.
    n = 3000
    st = iso_state(n, F, 5)
    e = make(n, F, CFG0, mode, E.FLAG_DIGEST)
    assert e.info().mode == mode and e.info().cells_per_axis == cells
    e.set_state(xyzv6=st)
    P = O.params(F, **CFG0)
    gp, gv = e.state_f64()
    for t in range(4):
        p, v = gp.copy(), gv.copy()
        _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
        e.step(1)
        bad = np.nonzero((e.digest() != d).any(1))[0]
        assert bad.size == 0, f"step {t}: {bad.size} boids with different neighbour sets, e.g. {bad[:5]}"
        gp, gv = e.state_f64()
        ep, ev = scaled_err(gp, gv, p, v, F, CFG0["vmax"])
        tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
        assert ep <= tol and ev <= tol, (t, ep, ev)
    say(f"{cells} cells per axis {E.MODE_NAMES[mode]} ok, rechecks={e.rechecks()}")


NO_CONE = dict(sep=(0.02, 1.0, 0.002), ali=(0.05, 1.5, 0.06), coh=(0.08, 2.0, 0.008), vmax=0.005, vmin=0.001)


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
@pytest.mark.parametrize("name", ["ordered", "no_cone"])
def test_rules_without_view_cone(mode, name):
    # a >= 1 skips the view cone (src/boid.cpp:29): the mixed kernel then
    # decides on distance alone (its cosine bounds are -inf / +inf), and a
    # NaN cosine (coincident points, zero velocity) must not change that.
    rules = {"ordered": ORDERED, "no_cone": NO_CONE}[name]
    n, F = 16384, 1.0
    P = O.params(F, **rules)
    e = make(n, F, rules, mode, E.FLAG_DIGEST)
    e.set_state(xyzv6=iso_state(n, F, 9, rules))
    gp, gv = e.state_f64()
    for t in range(4):
        p, v = gp.copy(), gv.copy()
        _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
        e.step(1)
        bad = np.nonzero((e.digest() != d).any(1))[0]
        assert bad.size == 0, f"step {t}: {bad.size} boids with different neighbour sets, e.g. {bad[:5]}"
        gp, gv = e.state_f64()
        ep, ev = scaled_err(gp, gv, p, v, F, rules["vmax"])
        tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
        assert ep <= tol and ev <= tol, (t, ep, ev)
    # a coincident pair and a boid at rest.  Without a view cone the pair
    # counts as separation neighbours, whose term (p_i - p_j) / |p_i - p_j|
    # is 0/0: the reference's velocity turns NaN there, and so must ours.
    st = np.array([[0.5, 0.5, 0.5, 0.003, 0, 0], [0.5, 0.5, 0.5, 0.003, 0, 0],
                   [0.51, 0.5, 0.5, 0.0, 0.0, 0.0], [0.995, 0.5, 0.5, 0.002, 0.001, 0]], np.float32)
    e = make(4, F, rules, mode, E.FLAG_DIGEST)
    e.set_state(xyzv6=st)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    # two updates: the second starts from NaN states, whose NaN distances make
    # them nobody's neighbours in the reference, so the NaN must not spread
    for t in range(2):
        e.step(1)
        _, d = O.step(P, p, v, digest=True)
        assert np.array_equal(e.digest(), d), t
        gp, gv = e.state_f64()
        assert np.array_equal(np.isnan(gp), np.isnan(p)) and np.array_equal(np.isnan(gv), np.isnan(v)), t
        ok = ~(np.isnan(p).any(1) | np.isnan(v).any(1))
        ep, ev = scaled_err(gp[ok], gv[ok], p[ok], v[ok], F, rules["vmax"])
        # the second update carries the first one's mixed-precision error
        tol = 1e-9 if t == 0 or mode == E.MODE_GRID_F64 else 1e-6
        assert ep < tol and ev < tol, (t, ep, ev)
    say(f"{name} {E.MODE_NAMES[mode]} ok")


CONE45 = dict(sep=(0.02, 0.25, 0.002), ali=(0.05, 0.25, 0.06), coh=(0.08, 0.25, 0.008), vmax=0.005, vmin=0.001)


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_owner_velocity_without_fp32_direction(mode):
    # An owner moving at 1e-20 has a well-defined FP64 direction while its
    # FP32 squared speed underflows: the mixed kernel may take no cone verdict
    # from its FP32 cosine.  With a 45-degree cone (a = 0.25) the boid ahead
    # is a cohesion and alignment neighbour, those behind and beside are not.
    F = 1.0
    st = np.array([[0.5, 0.5, 0.5, 1e-20, 0, 0], [0.53, 0.5, 0.5, 0.003, 0, 0],
                   [0.47, 0.5, 0.5, 0.003, 0, 0], [0.5, 0.53, 0.5, 0.003, 0, 0],
                   [0.5, 0.5, 0.515, 0.0, 0.0, -1e-20], [0.1, 0.1, 0.1, 0.003, 0, 0]], np.float64)
    e = make(len(st), F, CONE45, mode, E.FLAG_DIGEST)
    assert e.info().mode == mode
    e.set_state(pos=st[:, :3], vel=st[:, 3:])
    p, v = st[:, :3].copy(), st[:, 3:].copy()
    _, d = O.step(O.params(F, **CONE45), p, v, digest=True)
    e.step(1)
    gd = e.digest()
    assert (d[0, 0] & 0x1FFFFF) == 1 and (d[4, 0] & 0x1FFFFF) == 1, d[:, 0]   # the oracle's own verdicts
    assert np.array_equal(gd, d), (gd[:, 0], d[:, 0])
    gp, gv = e.state_f64()
    ep, ev = scaled_err(gp, gv, p, v, F, CONE45["vmax"])
    tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
    assert ep <= tol and ev <= tol, (ep, ev)


def test_grid_f64_ten_steps_continuous():
    n, F = 16384, 1.0
    st = iso_state(n, F, 1)
    e = make(n, F, CFG0, E.MODE_GRID_F64)
    e.set_state(xyzv6=st)
    P = O.params(F, **CFG0)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    for t in range(1, 11):
        O.step(P, p, v, enumeration=O.ENUM_CELLS)
        e.step(1)
        gp, gv = e.state_f64()
        ep, ev = scaled_err(gp, gv, p, v, F, CFG0["vmax"])
        tol = 1e-12 if t == 1 else 1e-9
        assert ep <= tol and ev <= tol, (t, ep, ev)


def test_grid_deterministic_and_resume_equals_continuous():
    n, F = 8192, 1.0
    st = iso_state(n, F, 5)
    a = make(n, F, CFG0, E.MODE_GRID_F64)
    a.set_state(xyzv6=st)
    a.step(10)
    pa, va = a.state_f64()
    b = make(n, F, CFG0, E.MODE_GRID_F64)
    b.set_state(xyzv6=st)
    b.step(4)
    p4, v4 = b.state_f64()
    c = make(n, F, CFG0, E.MODE_GRID_F64)
    c.set_state(pos=p4, vel=v4)          # sorted order depends only on the state
    c.step(6)
    pc, vc = c.state_f64()
    assert np.array_equal(pa, pc) and np.array_equal(va, vc)


def test_cfg2_full_size_one_step():
    # N = 2^20, F = 4 (density 16384): neighbour sets of one step, both modes
    n, F = 1 << 20, 4.0
    st = iso_state(n, F, 2)
    P = O.params(F, **CFG0)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    say("oracle step at N=2^20 ...")
    _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
    for mode in (E.MODE_GRID_F64, E.MODE_GRID_MIXED):
        e = make(n, F, CFG0, mode, E.FLAG_DIGEST)
        e.set_state(xyzv6=st)
        e.step(1)
        gd = e.digest()
        assert np.array_equal(gd, d), E.MODE_NAMES[mode]
        gp, gv = e.state_f64()
        ep, ev = scaled_err(gp, gv, p, v, F, CFG0["vmax"])
        assert ep <= (1e-12 if mode == E.MODE_GRID_F64 else 1e-6), ep
        assert ev <= (1e-12 if mode == E.MODE_GRID_F64 else 1e-6), ev
        # size-independent properties after more steps
        e.step(4)
        gp, gv = e.state_f64()
        s = np.linalg.norm(gv, axis=1)
        assert np.isfinite(gp).all() and (gp >= 0).all() and (gp <= F).all()
        assert (s >= CFG0["vmin"] * (1 - 1e-12)).all() and (s <= CFG0["vmax"] * (1 + 1e-12)).all()
        say(f"{E.MODE_NAMES[mode]} ok, rechecks={e.rechecks()}")


def test_every_workgroup_size_and_tile_depth():
    # The engine defaults to 64-thread workgroups; 128 and 256 (and every
    # tile depth) must give the same neighbour sets and states within the
    # bars, here on a state that cohesion has clustered for 200 updates.
    n, F = 1 << 16, (2.0 ** 16 / 16384.0) ** (1.0 / 3.0)
    e = make(n, F, CFG0, E.MODE_GRID_F64)
    e.set_state(xyzv6=iso_state(n, F, 5))
    e.step(200)
    p0, v0 = e.state_f64()
    e.close()
    P = O.params(F, **CFG0)
    p, v = p0.copy(), v0.copy()
    _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
    for mode in (E.MODE_GRID_F64, E.MODE_GRID_MIXED):
        tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
        for block in (64, 128, 256):
            for tile in (5, 6, 9, 12):
                g = B.Engine(B.make_params(n, F, **CFG0, mode=mode, flags=E.FLAG_DIGEST, block=block, tile=tile))
                info = g.info()
                assert (info.block, info.tile) == (block, tile)
                g.set_state(pos=p0, vel=v0)
                g.step(1)
                bad = np.nonzero((g.digest() != d).any(1))[0]
                assert bad.size == 0, (E.MODE_NAMES[mode], block, tile, bad.size)
                gp, gv = g.state_f64()
                ep, ev = scaled_err(gp, gv, p, v, F, CFG0["vmax"])
                assert ep <= tol and ev <= tol, (E.MODE_NAMES[mode], block, tile, ep, ev)
                g.close()
        say(f"{E.MODE_NAMES[mode]}: blocks 64/128/256 x tiles 5/6/9/12 ok")


def test_clustered_states_neighbour_sets_multi_pass_tiles():
    # Dense clusters: the headline workload after 300 updates (cells up to
    # ~60 boids) and the 64-blob clustered init at the 2^24-boid
    # configuration's density.  Tiles then exceed the LDS capacity and take
    # several passes (more with tile = 5); neighbour sets must stay exact.
    n, F = 1 << 20, 4.0
    e = make(n, F, CFG0, E.MODE_GRID_F64)
    e.set_state(xyzv6=iso_state(n, F, 2))
    e.step(300)
    evolved = e.state_f64()
    e.close()
    Fc = (n / (2.0 ** 24 / 1000.0)) ** (1.0 / 3.0)
    blobs = B.generate_init(E.INIT_CLUSTERED, n, Fc, CFG0["vmax"], CFG0["vmin"], seed=4)
    for label, Fs, (p0, v0) in (("evolved", F, evolved),
                                ("blobs", Fc, (blobs[:, :3].astype(np.float64), blobs[:, 3:].astype(np.float64)))):
        P = O.params(Fs, **CFG0)
        p, v = p0.copy(), v0.copy()
        say(f"oracle step ({label}) ...")
        _, d = O.step(P, p, v, digest=True, enumeration=O.ENUM_CELLS)
        for mode in (E.MODE_GRID_F64, E.MODE_GRID_MIXED):
            for tile in (5, 12):
                g = B.Engine(B.make_params(n, Fs, **CFG0, mode=mode, flags=E.FLAG_DIGEST, tile=tile))
                g.set_state(pos=p0, vel=v0)
                g.step(1)
                bad = np.nonzero((g.digest() != d).any(1))[0]
                assert bad.size == 0, (label, E.MODE_NAMES[mode], tile, bad.size)
                gp, gv = g.state_f64()
                ep, ev = scaled_err(gp, gv, p, v, Fs, CFG0["vmax"])
                tol = 1e-12 if mode == E.MODE_GRID_F64 else 1e-6
                assert ep <= tol and ev <= tol, (label, E.MODE_NAMES[mode], tile, ep, ev)
                say(f"{label} {E.MODE_NAMES[mode]} tile {tile} ok, rechecks {g.rechecks()}, "
                    f"max error {ep:.3g} F, {ev:.3g} vmax")
                g.close()


def test_cfg4_full_size_mixed_equals_fp64_neighbour_sets():
    # N = 2^24, F = 10, 64 Gaussian blobs (seed 4): the oracle would take
    # minutes here, so the size-independent check is that the mixed path's
    # neighbour sets equal the FP64 path's (both are exact) for 3 updates
    # from the same state, and that the state stays in the field with
    # vmin <= |v| <= vmax.
    n, F = 1 << 24, 10.0
    st = B.generate_init(E.INIT_CLUSTERED, n, F, CFG0["vmax"], CFG0["vmin"], seed=4)
    a = make(n, F, CFG0, E.MODE_GRID_F64, E.FLAG_DIGEST)
    b = make(n, F, CFG0, E.MODE_GRID_MIXED, E.FLAG_DIGEST)
    a.set_state(xyzv6=st)
    b.set_state(xyzv6=st)
    for t in range(3):
        a.step(1)
        b.step(1)
        da, db = a.digest(), b.digest()
        bad = np.nonzero((da != db).any(1))[0]
        assert bad.size == 0, (t, bad.size)
        pa, va = a.state_f64()
        b.set_state(pos=pa, vel=va)          # next step from the same state
    s = np.linalg.norm(va, axis=1)
    assert np.isfinite(pa).all() and (pa >= 0).all() and (pa <= F).all()
    assert (s >= CFG0["vmin"] * (1 - 1e-12)).all() and (s <= CFG0["vmax"] * (1 + 1e-12)).all()
    say(f"cfg4 ok, mixed rechecks {b.rechecks()}")


# ---------------------------------------------------------- edge cases ---
@pytest.mark.parametrize("mode", [E.MODE_ALLPAIRS_F32REF, E.MODE_ALLPAIRS_F64, E.MODE_GRID_F64, E.MODE_GRID_MIXED,
                                  E.MODE_GRID_F32REF])
def test_tiny_and_degenerate_inputs(mode):
    F = 1.0
    # n = 0
    e = make(0, F, CFG0, mode)
    e.set_state(xyzv6=np.zeros((0, 6), np.float32))
    e.step(3)
    # n = 1 crossing the wall; n = 4 with a coincident pair and a zero velocity
    st = np.array([[0.001, 0.5, 0.5, -0.004, 0.0, 0.0]], np.float32)
    e = make(1, F, CFG0, mode)
    e.set_state(xyzv6=st)
    e.step(1)
    p, v = e.state_f64()
    assert abs(p[0, 0] - (1.0 + (0.001 - 0.004))) < 1e-7 and abs(v[0, 0] + 0.004) < 1e-9
    st = np.array([[0.5, 0.5, 0.5, 0.003, 0, 0], [0.5, 0.5, 0.5, 0.003, 0, 0],
                   [0.51, 0.5, 0.5, 0.0, 0.0, 0.0], [0.995, 0.5, 0.5, 0.002, 0.001, 0]], np.float32)
    e = make(4, F, CFG0, mode, E.FLAG_DIGEST)
    e.set_state(xyzv6=st)
    e.step(1)
    P = O.params(F, **CFG0)
    if mode in (E.MODE_ALLPAIRS_F32REF, E.MODE_GRID_F32REF):
        p, v = st[:, :3].copy(), st[:, 3:].copy()
        _, d = O.step(P, p, v, digest=True)
        gp, gv = e.state_f32()
        assert np.array_equal(gp, p) and np.array_equal(gv, v)
    else:
        p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
        _, d = O.step(P, p, v, digest=True)
        gp, gv = e.state_f64()
        ep, ev = scaled_err(gp, gv, p, v, F, CFG0["vmax"])
        assert ep < 1e-9 and ev < 1e-9
    assert np.array_equal(e.digest(), d)


def test_small_field_falls_back_to_allpairs():
    # F / r_max < 3 cells per axis: the grid modes run all-pairs
    n, F = 500, 0.2
    st = iso_state(n, F, 6)
    e = make(n, F, CFG0, E.MODE_GRID_F64)
    assert e.info().mode == E.MODE_ALLPAIRS_F64
    e.set_state(xyzv6=st)
    e.step(3)
    P = O.params(F, **CFG0)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    for _ in range(3):
        O.step(P, p, v)
    gp, gv = e.state_f64()
    assert np.array_equal(gp, p) and np.array_equal(gv, v)


# -------------------------------------------------------------- metrics ---
@pytest.mark.parametrize("mode", [E.MODE_ALLPAIRS_F32REF, E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_metrics_match_oracle(mode):
    n, F = 6000, 0.72
    st = iso_state(n, F, 7)
    e = make(n, F, CFG0, mode)
    e.set_state(xyzv6=st)
    e.step(20)
    p, v = e.state_f64()
    for eps, ms in ((0.02, 2), (0.05, 1), (0.004, 2)):
        g = e.metrics(eps, ms)
        o = O.metrics(F, p, v, eps, ms)
        assert g["clusters"] == o["clusters"] and g["largest"] == o["largest"], (eps, g, o)
        assert abs(g["polarisation"] - o["polarisation"]) < 1e-12
        assert abs(g["mean_nn"] - o["mean_nn"]) < 1e-12 * F


def test_long_run_statistics_fp64_vs_mixed():
    # Chaotic runs diverge pointwise; their statistics must agree.  N = 2^20,
    # 600 steps on the GPU in both precisions; time-averaged over the last
    # 100 steps (every 20).  With these rules the flock stays disordered: its
    # polarisation is the random-direction value ~0.92/sqrt(N) (sd ~0.39/sqrt(N)),
    # so polarisation is compared against that noise floor; the ordered case
    # is test_long_run_polarisation_ordered_regime.
    n, F = 1 << 20, 4.0
    st = iso_state(n, F, 2)
    avg = {}
    for mode in (E.MODE_GRID_F64, E.MODE_GRID_MIXED):
        e = make(n, F, CFG0, mode)
        e.set_state(xyzv6=st)
        e.step(500)
        acc = []
        for _ in range(5):
            e.step(20)
            acc.append(e.metrics(0.02, 2))
        avg[mode] = {k: np.mean([a[k] for a in acc]) for k in acc[0]}
        say(E.MODE_NAMES[mode], avg[mode])
    a, b = avg[E.MODE_GRID_F64], avg[E.MODE_GRID_MIXED]
    for k in ("mean_nn", "clusters"):
        assert abs(a[k] - b[k]) <= 0.01 * abs(a[k]), (k, a[k], b[k])
    floor = 1.0 / np.sqrt(n)
    assert a["polarisation"] < 3 * floor and b["polarisation"] < 3 * floor
    assert abs(a["polarisation"] - b["polarisation"]) <= 3 * 0.39 * floor


# Strong alignment (radius 0.08, no view cone, weight 0.3) with a narrow
# speed range orders a flock within ~1000 updates (polarisation ~0.93).
ORDERED = dict(sep=(0.02, 0.75, 0.002), ali=(0.08, 1.0, 0.3), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.004)


def test_long_run_polarisation_ordered_regime():
    # Polarisation, mean nearest-neighbour distance and cluster count of
    # FP64 and mixed runs agree within 1% once the flock has ordered:
    # N = 65536 at density 16384, 3000 updates, averaged over the last 1500
    # (every 50) and over six seeds, each seed run in both precisions.
    # (Single chaotic runs differ by up to ~2% in cluster count.)
    n = 65536
    F = (n / 16384.0) ** (1.0 / 3.0)
    sums = {E.MODE_GRID_F64: {}, E.MODE_GRID_MIXED: {}}
    for seed in range(7, 13):
        st = B.generate_init(E.INIT_ISOTROPIC, n, F, ORDERED["vmax"], ORDERED["vmin"], seed=seed)
        for mode in (E.MODE_GRID_F64, E.MODE_GRID_MIXED):
            e = make(n, F, ORDERED, mode)
            e.set_state(xyzv6=st)
            e.step(1500)
            for _ in range(30):
                e.step(50)
                for k, v in e.metrics(0.02, 2).items():
                    sums[mode][k] = sums[mode].get(k, 0.0) + v
            e.close()
    a, b = sums[E.MODE_GRID_F64], sums[E.MODE_GRID_MIXED]
    say("fp64", {k: v / 180 for k, v in a.items()}, "mixed", {k: v / 180 for k, v in b.items()})
    assert a["polarisation"] / 180 > 0.8 and b["polarisation"] / 180 > 0.8
    for k in ("polarisation", "mean_nn", "clusters"):
        assert abs(a[k] - b[k]) <= 0.01 * abs(a[k]), (k, a[k] / 180, b[k] / 180)
