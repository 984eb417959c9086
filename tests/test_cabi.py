"""Host-side checks of libboids_engine.so that need no GPU: the library
loads, exports every entry point include/boids_engine.h declares, and its
host logic (init generators, init-file reader, exact decision thresholds)
agrees with the reference semantics."""
import ctypes
import ctypes.util
import math
import os
import re

import numpy as np
import pytest

import boidssimulatormpi_amd as B
from boidssimulatormpi_amd import engine as E
from oracle import oracle as O

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.acosf.restype = ctypes.c_float
libm.acosf.argtypes = [ctypes.c_float]
libm.acos.restype = ctypes.c_double
libm.acos.argtypes = [ctypes.c_double]


def test_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "boids_engine.h")).read()
    declared = set(re.findall(r"\b(bs_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(E.EXPORTS)
    L = E.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.bs_version().decode().startswith("boids_engine")


def test_generate_init_matches_reference_kinds(golden_dir):
    g = np.load(os.path.join(golden_dir, "ref_allpairs_ru_n1024.npz"))
    st = B.generate_init(E.INIT_RANDOM_UNIFORM, 1024, 1.0, 0.005, 0.001, seed=None)
    assert np.array_equal(st[:, :3], g["pos0"]) and np.array_equal(st[:, 3:], g["vel0"])
    g = np.load(os.path.join(golden_dir, "ref_grid_ru_seed1_n2048.npz"))
    st = B.generate_init(E.INIT_RANDOM_UNIFORM, 2048, 1.0, 0.005, 0.001, seed=1)
    assert np.array_equal(st[:, :3], g["pos0"]) and np.array_equal(st[:, 3:], g["vel0"])
    g = np.load(os.path.join(golden_dir, "ref_allpairs_testini_n1000.npz"))
    st = B.generate_init(E.INIT_TEST, 1000, 1.0, 0.003, 0.001, seed=None)
    assert np.array_equal(st[:, :3], g["pos0"]) and np.array_equal(st[:, 3:], g["vel0"])
    for seed in (0, 7):
        p, v = O.init_reference("test", 333, 2.0, 0.004, 0.002, seed=seed)
        st = B.generate_init(E.INIT_TEST, 333, 2.0, 0.004, 0.002, seed=seed)
        assert np.array_equal(st[:, :3], p) and np.array_equal(st[:, 3:], v)


@pytest.mark.parametrize("kind", [E.INIT_ISOTROPIC, E.INIT_CLUSTERED])
def test_generate_init_workloads(kind):
    n, F, vmax, vmin = 200000, 4.0, 0.005, 0.001
    st = B.generate_init(kind, n, F, vmax, vmin, seed=2)
    assert st.dtype == np.float32 and st.shape == (n, 6)
    p, v = st[:, :3].astype(np.float64), st[:, 3:].astype(np.float64)
    assert (p >= 0).all() and (p <= F).all()
    s = np.linalg.norm(v, axis=1)
    assert (s >= vmin * (1 - 1e-6)).all() and (s <= vmax * (1 + 1e-6)).all()
    again = B.generate_init(kind, n, F, vmax, vmin, seed=2)
    assert np.array_equal(st, again)
    assert not np.array_equal(st, B.generate_init(kind, n, F, vmax, vmin, seed=3))
    if kind == E.INIT_ISOTROPIC:
        d = v / s[:, None]
        assert np.abs(d.mean(0)).max() < 0.01                # isotropic directions
        assert abs(np.mean(d[:, 2] ** 2) - 1 / 3) < 0.01
        assert np.abs(p.mean(0) - F / 2).max() < 0.02
    else:
        # uneven occupancy: the densest 1% of cells hold many times the mean
        G = 50
        c = np.minimum((p / F * G).astype(int), G - 1)
        occ = np.bincount(c[:, 0] + G * (c[:, 1] + G * c[:, 2]), minlength=G ** 3)
        assert np.sort(occ)[-G ** 3 // 100:].mean() > 5 * occ.mean()


def test_read_init_file(tmp_path):
    st = B.generate_init(E.INIT_ISOTROPIC, 100, 1.0, 0.005, 0.001, seed=9)
    f = tmp_path / "init.bin"
    st.tofile(f)
    out = np.zeros((100, 6), np.float32)
    rc = E.lib().bs_read_init_file(str(f).encode(), 100, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert rc == 0 and np.array_equal(out, st)
    rc = E.lib().bs_read_init_file(str(f).encode(), 101, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert rc == E.lib().bs_read_init_file(b"/nonexistent", 1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) < 0


def f32_prev(x):
    return np.nextafter(np.float32(x), np.float32(-np.inf))


@pytest.mark.parametrize("rules", [
    dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008)),
    dict(sep=(0.1, 0.9, 0.0001), ali=(0.1, 0.7, 0.1), coh=(0.1, 0.8, 0.1)),
    dict(sep=(0.01, 0.5, 0.002), ali=(0.05, 1 / 3, 0.06), coh=(0.05, 0.5, 0.008)),
    dict(sep=(0.05, 1.0, 0.0001), ali=(0.05, 0.0, 0.1), coh=(0.05, -0.1, 0.1)),
])
def test_exact_thresholds(rules):
    # float: s* smallest with (double)sqrtf(s*) >= r; c* smallest with acosf(c*) <= pi*a
    radii = [rules["sep"][0], rules["ali"][0], rules["coh"][0]]
    angles = [rules["sep"][1], rules["ali"][1], rules["coh"][1]]
    t32 = B.engine.debug_thresholds(B.make_params(10, 1.0, **rules, mode=E.MODE_ALLPAIRS_F32REF))
    t64 = B.engine.debug_thresholds(B.make_params(10, 1.0, **rules, mode=E.MODE_GRID_F64))
    for k in range(3):
        r, a = radii[k], angles[k]
        s = np.float32(t32[k])
        assert float(np.sqrt(s)) >= r and float(np.sqrt(f32_prev(s))) < r
        s = t64[k]
        assert math.sqrt(s) >= r and math.sqrt(np.nextafter(s, 0)) < r
        if a >= 1:
            assert t32[3 + k] == -2.0 and t64[3 + k] == -2.0
            continue
        th = math.pi * a
        c = t32[3 + k]
        if c == 2.0:
            assert not (float(libm.acosf(1.0)) <= th)
        else:
            c = np.float32(c)
            assert float(libm.acosf(c)) <= th and not (float(libm.acosf(f32_prev(c))) <= th)
            # a window of floats each side agrees with the threshold form
            for o in range(1, 2000, 37):
                hi = np.float32(c)
                lo = np.float32(c)
                for _ in range(o):
                    hi = np.nextafter(hi, np.float32(2)); lo = np.nextafter(lo, np.float32(-2))
                if hi <= 1:
                    assert float(libm.acosf(hi)) <= th
                assert not (float(libm.acosf(lo)) <= th)
        c = t64[3 + k]
        if c == 2.0:
            assert not (libm.acos(1.0) <= th)
        else:
            assert libm.acos(c) <= th and not (libm.acos(np.nextafter(c, -2.0)) <= th)
