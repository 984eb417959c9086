"""Pins the CPU oracle (oracle/) before anything is checked against it.

- float instantiation vs golden fixtures written by the reference's own
  compiled engines (tests/golden/make_golden.py): bit-for-bit;
- drand48 restatement vs the C library's drand48;
- FP64 instantiation vs closed forms and special cases a plausible slip
  (sign of the shift, dropped term, wrong index, cone direction) would fail;
- cell enumeration vs all-pairs, bit-for-bit, in both precisions.
"""
import ctypes
import ctypes.util
import hashlib
import math
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O

CFG0 = dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001)


def params_from(rules):
    F, sr, sa, sw, ar, aa, aw, cr, ca, cw, vmax, vmin = rules
    return O.params(F, (sr, sa, sw), (ar, aa, aw), (cr, ca, cw), vmax, vmin)


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


# ---------------------------------------------------------------- drand48 ---
def test_drand48_matches_libc():
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (0, 1, 2, 12345, -7, 2**31 - 1):
        libc.srand48(seed)
        ref = np.array([libc.drand48() for _ in range(200)])
        assert np.array_equal(ref, O.drand48_stream(seed, 200)), seed


def test_drand48_unseeded_state_is_zero():
    # X_0 = 0: the first value is c / 2^48 = 11 / 2^48
    v = O.drand48_stream(None, 2)
    assert v[0] == 11.0 / 2**48
    assert v[1] == ((0x5DEECE66D * 11 + 11) % 2**48) / 2**48


# -------------------------------------------------- float vs the reference ---
def test_init_matches_reference(golden_dir):
    g = load(golden_dir, "ref_allpairs_ru_n1024.npz")
    p, v = O.init_reference("random_uniform", 1024, 1.0, 0.005, 0.001, seed=None)
    assert np.array_equal(p, g["pos0"]) and np.array_equal(v, g["vel0"])
    g = load(golden_dir, "ref_grid_ru_seed1_n2048.npz")
    p, v = O.init_reference("random_uniform", 2048, 1.0, 0.005, 0.001, seed=1)
    assert np.array_equal(p, g["pos0"]) and np.array_equal(v, g["vel0"])
    g = load(golden_dir, "ref_allpairs_testini_n1000.npz")
    p, v = O.init_reference("test", 1000, 1.0, 0.003, 0.001, seed=None)
    assert np.array_equal(p, g["pos0"]) and np.array_equal(v, g["vel0"])


@pytest.mark.parametrize("enum", [O.ENUM_ALLPAIRS, O.ENUM_CELLS])
def test_f32_allpairs_reference(golden_dir, enum):
    g = load(golden_dir, "ref_allpairs_ru_n1024.npz")
    P = params_from(g["rules"])
    p, v = g["pos0"].copy(), g["vel0"].copy()
    w = np.float32([P.coh_w] * 3 + [P.sep_w] * 3 + [P.ali_w] * 3)
    for t in range(1, 101):
        f, _ = O.step(P, p, v, forces=True, enumeration=enum)
        if t in (1, 100):
            assert np.array_equal(p, g[f"pos{t}"]) and np.array_equal(v, g[f"vel{t}"]), t
            assert np.array_equal(f * w, g[f"force{t}"]), t   # get_force = (float)w * dv


@pytest.mark.parametrize("enum", [O.ENUM_ALLPAIRS, O.ENUM_CELLS])
def test_f32_file_init_reference(golden_dir, enum):
    g = load(golden_dir, "ref_allpairs_file_n1024.npz")
    P = params_from(g["rules"])
    init = g["init"]
    p, v = init[:, :3].copy(), init[:, 3:].copy()
    w = np.float32([P.coh_w] * 3 + [P.sep_w] * 3 + [P.ali_w] * 3)
    for t in range(1, 101):
        f, _ = O.step(P, p, v, forces=True, enumeration=enum)
        if t in (1, 10, 100):
            assert np.array_equal(p, g[f"pos{t}"]) and np.array_equal(v, g[f"vel{t}"]), t
        if t in (1, 100):
            assert np.array_equal(f * w, g[f"force{t}"]), t


def test_f32_grid_reference(golden_dir):
    # boidsim_grid == boidsim on one init (GRID_NUM >= 2); both enumerations
    g = load(golden_dir, "ref_grid_ru_seed1_n2048.npz")
    P = params_from(g["rules"])
    for enum in (O.ENUM_ALLPAIRS, O.ENUM_CELLS):
        p, v = g["pos0"].copy(), g["vel0"].copy()
        for _ in range(50):
            O.step(P, p, v, enumeration=enum)
        assert np.array_equal(p, g["pos50"]) and np.array_equal(v, g["vel50"])


def test_f32_distance_only_reference(golden_dir):
    g = load(golden_dir, "ref_grid_distonly_n2048.npz")
    P = params_from(g["rules"])
    p, v = g["pos0"].copy(), g["vel0"].copy()
    for _ in range(20):
        O.step(P, p, v, enumeration=O.ENUM_CELLS)
    assert np.array_equal(p, g["pos20"]) and np.array_equal(v, g["vel20"])


def ptcl(frames, n, T, F):
    hdr = struct.pack("<4sBBIIIIIffffff", b"PTCL", 0, 2, 50, n, T, 0, 30, 0.0, F, 0.0, F, 0.0, F)
    return hdr + b"".join(f.astype("<f4").tobytes() for f in frames)


def test_f32_testini_ptcl(golden_dir):
    # the whole boidsim output for settings/test.ini, byte for byte
    g = load(golden_dir, "ref_allpairs_testini_n1000.npz")
    P = params_from(g["rules"])
    p, v = g["pos0"].copy(), g["vel0"].copy()
    frames, vframes = [], []
    for _ in range(30):
        frames.append(p.copy()); vframes.append(v.copy())
        O.step(P, p, v)
    assert hashlib.sha256(ptcl(frames, 1000, 30, 1.0)).digest() == g["ptcl_sha256"].tobytes()
    assert hashlib.sha256(ptcl(vframes, 1000, 30, 1.0)).digest() == g["vptcl_sha256"].tobytes()
    assert np.array_equal(p, g["pos30"]) and np.array_equal(v, g["vel30"])


@pytest.mark.parametrize("enum", [O.ENUM_ALLPAIRS, O.ENUM_CELLS])
def test_f32_grid_testini_ptcl(golden_dir, enum):
    # the whole boidsim_grid output for settings/test.ini (its own init: the
    # "test" kind after srand48(12345)), byte for byte
    g = load(golden_dir, "ref_grid_testini_n1000.npz")
    P = params_from(g["rules"])
    p, v = O.init_reference("test", 1000, 1.0, 0.003, 0.001, seed=12345)
    assert np.array_equal(p, g["pos0"]) and np.array_equal(v, g["vel0"])
    frames, vframes = [], []
    for _ in range(30):
        frames.append(p.copy()); vframes.append(v.copy())
        O.step(P, p, v, enumeration=enum)
    assert hashlib.sha256(ptcl(frames, 1000, 30, 1.0)).digest() == g["ptcl_sha256"].tobytes()
    assert hashlib.sha256(ptcl(vframes, 1000, 30, 1.0)).digest() == g["vptcl_sha256"].tobytes()
    assert np.array_equal(p, g["pos30"]) and np.array_equal(v, g["vel30"])


def test_f32_grid_cfg1_full_size_ptcl(golden_dir):
    # boidsim_grid at the full cfg1 size (N = 16384, F = 1, cfg0 rules, its
    # random_uniform init after srand48(1)): the oracle's cell enumeration
    # reproduces the reference grid engine's .ptcl digests over 200 updates
    g = load(golden_dir, "ref_grid_cfg1_n16384.npz")
    P = params_from(g["rules"])
    n, T = 16384, int(g["steps"])
    p, v = O.init_reference("random_uniform", n, 1.0, 0.005, 0.001, seed=1)
    assert hashlib.sha256(np.concatenate([p, v], 1).astype("<f4").tobytes()).digest() == g["init_sha256"].tobytes()
    hp = hashlib.sha256(ptcl([], n, T, 1.0))          # the header alone
    hv = hashlib.sha256(ptcl([], n, T, 1.0))
    for _ in range(T):
        hp.update(p.astype("<f4").tobytes()); hv.update(v.astype("<f4").tobytes())
        O.step(P, p, v, enumeration=O.ENUM_CELLS)
    assert hp.digest() == g["ptcl_sha256"].tobytes() and hv.digest() == g["vptcl_sha256"].tobytes()
    assert hashlib.sha256(np.concatenate([p, v], 1).astype("<f4").tobytes()).digest() == g["final_sha256"].tobytes()


# --------------------------------------------------- FP64: closed forms ---
def f64(pos, vel):
    return np.array(pos, np.float64).reshape(-1, 3), np.array(vel, np.float64).reshape(-1, 3)


def test_f64_single_boid_free_flight_clamp_wrap():
    P = O.params(1.0, **CFG0)
    # too fast -> renormalised to vmax, then moves; crosses x = 0 and wraps to F + x
    p, v = f64([0.001, 0.5, 0.5], [-0.006, 0.0, 0.0])
    f, d = O.step(P, p, v, forces=True, digest=True)
    assert v[0, 0] == -0.005 and v[0, 1] == 0 and v[0, 2] == 0
    assert p[0, 0] == 1.0 + (0.001 - 0.005) and p[0, 1] == 0.5
    assert not f.any() and not d.any()
    # too slow -> vmin; exactly at vmax -> untouched
    p, v = f64([0.5, 0.5, 0.5], [0.0, 0.0002, 0.0])
    O.step(P, p, v)
    assert v[0, 1] == 0.001
    p, v = f64([0.5, 0.5, 0.5], [0.0, 0.0, 0.005])
    O.step(P, p, v)
    assert v[0, 2] == 0.005 and p[0, 2] == 0.505
    # zero velocity stays zero (norm 0 skips both clamps)
    p, v = f64([0.5, 0.5, 0.5], [0.0, 0.0, 0.0])
    O.step(P, p, v)
    assert not v.any() and p[0, 0] == 0.5


def test_f64_wrap_upper_boundary():
    P = O.params(1.0, **CFG0)
    p, v = f64([0.999, 0.2, 0.3], [0.004, 0.0, 0.0])
    O.step(P, p, v)
    assert p[0, 0] == (0.999 + 0.004) - 1.0
    p, v = f64([0.996, 0.2, 0.3], [0.004, 0.0, 0.0])   # lands on exactly F? stays (p == F is kept)
    O.step(P, p, v)
    assert p[0, 0] == 0.996 + 0.004


def two_boids(P, pi, vi, pj, vj):
    p, v = f64([*pi, *pj], [*vi, *vj])
    f, d = O.step(P, p, v, forces=True, digest=True)
    return p, v, f, d


def counts(d):
    c = int(d[0])
    return c & (2**21 - 1), (c >> 21) & (2**21 - 1), (c >> 42) & (2**21 - 1)


def test_f64_view_cone_ahead_behind_side():
    P = O.params(1.0, **CFG0)    # 270 degree view: blind cone of 45 deg half-angle behind
    vi = (0.003, 0.0, 0.0)
    # j ahead at 0.01: inside sep, ali, coh radii and cone
    _, _, f, d = two_boids(P, (0.5, 0.5, 0.5), vi, (0.51, 0.5, 0.5), (0.0, 0.003, 0.0))
    assert counts(d[0]) == (1, 1, 1)
    # coh: q - p_i; sep: unit vector away from j; ali: v_j - v_i
    assert np.allclose(f[0, 0:3], [0.01, 0, 0], atol=1e-15)
    assert np.allclose(f[0, 3:6], [-1, 0, 0], atol=1e-15)
    assert np.allclose(f[0, 6:9], [-0.003, 0.003, 0], atol=1e-15)
    # j straight behind: acos(-1) = pi > 0.75 pi -> unseen by i; i is ahead of j?
    _, _, f, d = two_boids(P, (0.5, 0.5, 0.5), vi, (0.49, 0.5, 0.5), vi)
    assert counts(d[0]) == (0, 0, 0) and counts(d[1]) == (1, 1, 1)
    # 100 deg off the heading (< 135): seen; 150 deg: unseen
    for ang, seen in ((100, 1), (150, 0)):
        a = math.radians(ang)
        _, _, _, d = two_boids(P, (0.5, 0.5, 0.5), vi, (0.5 + 0.01 * math.cos(a), 0.5 + 0.01 * math.sin(a), 0.5), vi)
        assert counts(d[0]) == (seen, seen, seen), ang


def test_f64_distance_only_and_radii():
    P = O.params(1.0, sep=(0.02, 1.0, 0.002), ali=(0.05, 1.0, 0.06), coh=(0.08, 1.0, 0.008))
    vi = (0.003, 0.0, 0.0)
    # behind but a >= 1: seen; distances select rules: 0.03 -> ali+coh, 0.06 -> coh only
    for dist, expect in ((0.01, (1, 1, 1)), (0.03, (1, 0, 1)), (0.06, (1, 0, 0)), (0.09, (0, 0, 0))):
        _, _, _, d = two_boids(P, (0.5, 0.5, 0.5), vi, (0.5 - dist, 0.5, 0.5), vi)
        assert counts(d[0]) == expect, dist


def test_f64_minimum_image_across_boundary():
    P = O.params(1.0, **CFG0)
    # i near x = 0, j near x = F: the shifted q = x_j - F lies 0.02 behind... put j ahead across the wrap
    _, _, f, d = two_boids(P, (0.995, 0.5, 0.5), (0.003, 0.0, 0.0), (0.005, 0.5, 0.5), (0.003, 0.0, 0.0))
    assert counts(d[0])[0] == 1
    assert abs(f[0, 0] - ((0.005 + 1.0) - 0.995)) < 1e-15          # cohesion towards +x across the wrap
    # and j behind across the wrap for i near 0 is unseen (blind cone)
    _, _, _, d = two_boids(P, (0.005, 0.5, 0.5), (0.003, 0.0, 0.0), (0.995, 0.5, 0.5), (0.003, 0.0, 0.0))
    assert counts(d[0]) == (0, 0, 0)


def test_f64_zero_velocity_and_coincident():
    P = O.params(1.0, **CFG0)
    # zero own velocity -> cosine NaN -> no neighbour (cone rules)
    _, _, _, d = two_boids(P, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), (0.51, 0.5, 0.5), (0.003, 0, 0))
    assert counts(d[0]) == (0, 0, 0)
    # coincident distinct boids -> NaN cosine -> rejected
    _, _, _, d = two_boids(P, (0.5, 0.5, 0.5), (0.003, 0.0, 0.0), (0.5, 0.5, 0.5), (0.003, 0, 0))
    assert counts(d[0]) == (0, 0, 0)


def test_f64_weights_zero_is_ballistic():
    P = O.params(1.0, sep=(0.02, 0.75, 0.0), ali=(0.05, 0.75, 0.0), coh=(0.08, 0.75, 0.0))
    rng = np.random.default_rng(3)
    p = rng.random((300, 3)) * 0.2 + 0.3     # away from the walls: no wrap
    v = rng.normal(size=(300, 3))
    v *= (0.001 + 0.004 * rng.random(300))[:, None] / np.linalg.norm(v, axis=1)[:, None]
    p0, v0 = p.copy(), v.copy()
    O.step(P, p, v)
    assert np.array_equal(v, v0)
    assert np.array_equal(p, p0 + v0)


def test_f64_constant_speed():
    P = O.params(1.0, **{**CFG0, "vmax": 0.003, "vmin": 0.003})
    rng = np.random.default_rng(4)
    p = rng.random((500, 3)).astype(np.float64)
    v = rng.normal(size=(500, 3)) * 0.003
    for _ in range(5):
        O.step(P, p, v)
    assert np.allclose(np.linalg.norm(v, axis=1), 0.003, rtol=1e-14)


# ------------------------------------------- enumeration equivalences ---
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("rules", [CFG0,
                                   dict(sep=(0.01, 0.5, 0.002), ali=(0.05, 1 / 3, 0.06), coh=(0.05, 0.5, 0.008), vmax=0.005, vmin=0.001),
                                   dict(sep=(0.05, 1.0, 0.0001), ali=(0.05, 1.0, 0.1), coh=(0.05, 1.0, 0.1), vmax=0.003, vmin=0.001)])
def test_cells_equal_allpairs(dtype, rules):
    rng = np.random.default_rng(11)
    n, F = 4096, 1.0
    p = (rng.random((n, 3)) * F).astype(dtype)
    v = rng.normal(size=(n, 3))
    v = (v * ((0.001 + 0.004 * rng.random(n)) / np.linalg.norm(v, axis=1))[:, None]).astype(dtype)
    P = O.params(F, **rules)
    assert O.cells_per_axis(P) >= 3
    pa, va, pb, vb = p.copy(), v.copy(), p.copy(), v.copy()
    for _ in range(5):
        fa, da = O.step(P, pa, va, forces=True, digest=True, enumeration=O.ENUM_ALLPAIRS)
        fb, db = O.step(P, pb, vb, forces=True, digest=True, enumeration=O.ENUM_CELLS)
        assert np.array_equal(da, db)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb) and np.array_equal(fa, fb)


def test_digest_is_a_neighbour_set_hash():
    # digest = (counts, sum of mix64(j)) per rule: check by brute force on tiny input
    P = O.params(1.0, **CFG0)
    rng = np.random.default_rng(5)
    n = 60
    p = rng.random((n, 3)) * 0.12 + 0.4
    v = rng.normal(size=(n, 3)) * 0.003
    pp, vv = p.copy(), v.copy()
    _, d = O.step(P, pp, vv, digest=True)
    for i in range(n):
        nb = {0: [], 1: [], 2: []}
        for j in range(n):
            if j == i:
                continue
            e = p[i] - p[j]
            dist = math.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2])
            g = -e
            c = (v[i, 0] * g[0] + v[i, 1] * g[1] + v[i, 2] * g[2]) / (
                math.sqrt((v[i, 0] ** 2 + v[i, 1] ** 2) + v[i, 2] ** 2) * dist)
            ok = math.acos(max(-1.0, min(1.0, c))) <= math.pi * 0.75 if abs(c) <= 1 else False
            for k, r in enumerate((0.08, 0.02, 0.05)):     # coh, sep, ali
                if dist < r and ok:
                    nb[k].append(j)
        assert counts(d[i]) == (len(nb[0]), len(nb[1]), len(nb[2]))
        for k in range(3):
            assert int(d[i, 1 + k]) == sum(O.mix64(j) for j in nb[k]) % 2**64


# ------------------------------------------------------------- metrics ---
def test_metrics_closed_forms():
    # polarisation: aligned -> 1, opposed pairs -> 0; zero velocities add nothing
    pos = np.array([[0.1, 0.1, 0.1], [0.5, 0.5, 0.5], [0.9, 0.2, 0.4], [0.3, 0.7, 0.6]])
    vel = np.tile([0.0, 0.003, 0.0], (4, 1))
    assert abs(O.metrics(1.0, pos, vel, 0.01)["polarisation"] - 1.0) < 1e-15
    vel[1] *= -1; vel[3] *= -1
    assert O.metrics(1.0, pos, vel, 0.01)["polarisation"] < 1e-15
    vel[:] = 0
    assert O.metrics(1.0, pos, vel, 0.01)["polarisation"] == 0.0
    # nearest neighbour on a cubic lattice of spacing 0.1 (periodic: every
    # point has six neighbours at 0.1, also across the boundary)
    g = (np.arange(10) + 0.5) / 10
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    m = O.metrics(1.0, lat, np.zeros_like(lat), 0.1001, 1)
    assert abs(m["mean_nn"] - 0.1) < 1e-12
    assert m["clusters"] == 1 and m["largest"] == 1000          # eps > spacing: one component
    m = O.metrics(1.0, lat, np.zeros_like(lat), 0.0999, 1)
    assert m["clusters"] == 1000 and m["largest"] == 1
    assert O.metrics(1.0, lat, np.zeros_like(lat), 0.0999, 2)["clusters"] == 0
    # periodic distance across the wall, and chains
    pos = np.array([[0.005, 0.5, 0.5], [0.995, 0.5, 0.5], [0.5, 0.5, 0.5], [0.51, 0.5, 0.5], [0.52, 0.5, 0.5]])
    m = O.metrics(1.0, pos, np.ones_like(pos), 0.0105, 2)
    assert m["clusters"] == 2 and m["largest"] == 3
    assert abs(m["mean_nn"] - 0.01) < 1e-12
