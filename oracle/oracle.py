"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  The restated reference
functions are cited in boids_oracle.h / boids_oracle_step.inc.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

ENUM_ALLPAIRS = 0
ENUM_CELLS = 1


class Params(ctypes.Structure):
    """bo_params: field, (r, a, w) for separation, alignment, cohesion, vmax, vmin."""

    _fields_ = [(k, ctypes.c_double) for k in (
        "field", "sep_r", "sep_a", "sep_w", "ali_r", "ali_a", "ali_w",
        "coh_r", "coh_a", "coh_w", "vmax", "vmin")]


class Rng(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint64)]


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        for name, real in (("bo_step_f32", ctypes.c_float), ("bo_step_f64", ctypes.c_double)):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [P(Params), ctypes.c_uint32, P(real), P(real), P(real),
                          P(ctypes.c_uint64), ctypes.c_int, ctypes.c_int]
        for name, real in (("bo_time_step_sample_f32", ctypes.c_float), ("bo_time_step_sample_f64", ctypes.c_double)):
            f = getattr(L, name)
            f.restype = ctypes.c_long
            f.argtypes = [P(Params), ctypes.c_uint32, P(real), P(real), ctypes.c_uint32, ctypes.c_int,
                          P(ctypes.c_double)]
        L.bo_cells_per_axis.restype = ctypes.c_int
        L.bo_cells_per_axis.argtypes = [P(Params)]
        L.bo_mix64.restype = ctypes.c_uint64
        L.bo_mix64.argtypes = [ctypes.c_uint64]
        L.bo_srand48.argtypes = [P(Rng), ctypes.c_long]
        L.bo_rng_unseeded.argtypes = [P(Rng)]
        L.bo_drand48.restype = ctypes.c_double
        L.bo_drand48.argtypes = [P(Rng)]
        L.bo_init_reference.restype = ctypes.c_int
        L.bo_init_reference.argtypes = [ctypes.c_int, P(Rng), ctypes.c_uint32, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, P(ctypes.c_float),
                                        P(ctypes.c_float)]
        L.bo_metrics.restype = ctypes.c_int
        L.bo_metrics.argtypes = [ctypes.c_uint32, ctypes.c_double, P(ctypes.c_double), P(ctypes.c_double),
                                 ctypes.c_double, ctypes.c_uint32, ctypes.c_int, P(MetricsOut)]
        _lib = L
    return _lib


class MetricsOut(ctypes.Structure):
    _fields_ = [("polarisation", ctypes.c_double), ("mean_nn", ctypes.c_double),
                ("clusters", ctypes.c_uint64), ("largest", ctypes.c_uint64)]


def metrics(field, pos, vel, eps, min_size=2, nthreads=0) -> dict:
    """Brute-force flock metrics (see bo_metrics in boids_oracle.h)."""
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    out = MetricsOut()
    rc = lib().bo_metrics(pos.shape[0], field, _ptr(pos, ctypes.c_double), _ptr(vel, ctypes.c_double),
                          eps, min_size, nthreads, ctypes.byref(out))
    if rc != 0:
        raise RuntimeError("oracle metrics failed")
    return {k: getattr(out, k) for k, _ in out._fields_}


def params(field, sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008),
           vmax=0.005, vmin=0.001) -> Params:
    return Params(field, *sep, *ali, *coh, vmax, vmin)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


def step(p: Params, pos, vel, *, forces=False, digest=False, enumeration=ENUM_ALLPAIRS,
         nthreads=0):
    """One update() in place on pos/vel (float32 -> reference-faithful f32
    instantiation, float64 -> f64 instantiation).  Returns (forces, digest).
    nthreads: OpenMP threads (0: every host thread)."""
    assert pos.dtype == vel.dtype and pos.flags.c_contiguous and vel.flags.c_contiguous
    n = pos.shape[0]
    assert pos.shape == (n, 3) and vel.shape == (n, 3)
    ct = ctypes.c_float if pos.dtype == np.float32 else ctypes.c_double
    fn = lib().bo_step_f32 if pos.dtype == np.float32 else lib().bo_step_f64
    f = np.zeros((n, 9), dtype=pos.dtype) if forces else None
    d = np.zeros((n, 4), dtype=np.uint64) if digest else None
    rc = fn(ctypes.byref(p), n, _ptr(pos, ct), _ptr(vel, ct), _ptr(f, ct),
            _ptr(d, ctypes.c_uint64), enumeration, nthreads)
    if rc != 0:
        raise RuntimeError("oracle step failed")
    return f, d


def time_step_sample(p: Params, pos, vel, stride, nthreads=0):
    """Timed parts of one cell-enumeration update without changing the state:
    (sampled owners, build s, phase A of every stride-th owner s, phase B s)."""
    pos = np.ascontiguousarray(pos)
    vel = np.ascontiguousarray(vel, dtype=pos.dtype)
    ct = ctypes.c_float if pos.dtype == np.float32 else ctypes.c_double
    fn = lib().bo_time_step_sample_f32 if pos.dtype == np.float32 else lib().bo_time_step_sample_f64
    sec = np.zeros(3, np.float64)
    k = fn(ctypes.byref(p), pos.shape[0], _ptr(pos, ct), _ptr(vel, ct), stride, nthreads, _ptr(sec, ctypes.c_double))
    if k < 0:
        raise RuntimeError("oracle timing sample failed")
    return int(k), float(sec[0]), float(sec[1]), float(sec[2])


def cells_per_axis(p: Params) -> int:
    return lib().bo_cells_per_axis(ctypes.byref(p))


def mix64(j: int) -> int:
    return lib().bo_mix64(j)


def init_reference(kind: str, n: int, field: float, vmax: float, vmin: float, seed=None):
    """Reference init ("test" | "random_uniform").  seed None reproduces the
    all-pairs engine (drand48 never seeded); an int reproduces the grid
    engine's srand48(seed)."""
    r = Rng()
    if seed is None:
        lib().bo_rng_unseeded(ctypes.byref(r))
    else:
        lib().bo_srand48(ctypes.byref(r), seed)
    pos = np.zeros((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    k = {"test": 0, "random_uniform": 1}[kind]
    rc = lib().bo_init_reference(k, ctypes.byref(r), n, field, vmax, vmin,
                                 _ptr(pos, ctypes.c_float), _ptr(vel, ctypes.c_float))
    assert rc == 0
    return pos, vel


def drand48_stream(seed, count):
    r = Rng()
    if seed is None:
        lib().bo_rng_unseeded(ctypes.byref(r))
    else:
        lib().bo_srand48(ctypes.byref(r), seed)
    return np.array([lib().bo_drand48(ctypes.byref(r)) for _ in range(count)])


def read_ref_dump(path: str, n: int):
    """Parse a ref_harness dump: {t: (pos, vel, forces9)} as float32 arrays."""
    raw = open(path, "rb").read()
    rec = 4 + 4 * 15 * n
    assert len(raw) % rec == 0, "truncated dump"
    out = {}
    for k in range(len(raw) // rec):
        b = raw[k * rec:(k + 1) * rec]
        t = int(np.frombuffer(b[:4], np.uint32)[0])
        a = np.frombuffer(b[4:], np.float32)
        out[t] = (a[:3 * n].reshape(n, 3).copy(), a[3 * n:6 * n].reshape(n, 3).copy(),
                  a[6 * n:].reshape(n, 9).copy())
    return out
