"""ctypes binding of libboids_engine.so (C ABI in include/boids_engine.h).

`Engine` is a thin handle over the C ABI (bulk state in/out, step, timing).
`BoidSimulation` mirrors the reference class of the same name
(src/boid_simulation.h:15-71): setup / init / update / get / get_v /
get_force with the same argument order and meaning, on top of `Engine`.
There is no CPU fallback: if the library is missing, importing this module
raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# BOIDS_ENGINE_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("BOIDS_ENGINE_LIB") or os.path.join(PKG_DIR, "libboids_engine.so")

MODE_ALLPAIRS_F32REF = 0
MODE_ALLPAIRS_F64 = 1
MODE_GRID_F64 = 2
MODE_GRID_MIXED = 3
MODE_GRID_F32REF = 4
MODE_NAMES = {MODE_ALLPAIRS_F32REF: "allpairs_f32ref", MODE_ALLPAIRS_F64: "allpairs_f64",
              MODE_GRID_F64: "grid_f64", MODE_GRID_MIXED: "grid_mixed", MODE_GRID_F32REF: "grid_f32ref"}

INIT_TEST = 0
INIT_RANDOM_UNIFORM = 1
INIT_ISOTROPIC = 2
INIT_CLUSTERED = 3

FLAG_FORCES = 1
FLAG_DIGEST = 2
FLAG_TIMING = 4
FLAG_GRAPH = 8

K_PAIR, K_SCAN, K_SCATTER, K_GATHER, K_OTHER, K_STEP = range(6)
K_NAMES = ["pair", "scan", "scatter", "gather", "other", "step"]


class Params(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("field", ctypes.c_double),
                ("sep_r", ctypes.c_double), ("sep_a", ctypes.c_double), ("sep_w", ctypes.c_double),
                ("ali_r", ctypes.c_double), ("ali_a", ctypes.c_double), ("ali_w", ctypes.c_double),
                ("coh_r", ctypes.c_double), ("coh_a", ctypes.c_double), ("coh_w", ctypes.c_double),
                ("vmax", ctypes.c_double), ("vmin", ctypes.c_double),
                ("mode", ctypes.c_int32), ("cells_per_axis", ctypes.c_int32),
                ("device", ctypes.c_int32), ("block", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("fill", ctypes.c_int32)]


class Info(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("cells_per_axis", ctypes.c_int32),
                ("cell_edge", ctypes.c_double), ("device_bytes", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("block", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("fill", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class FlockInfo(ctypes.Structure):
    _fields_ = [("boid_base", ctypes.c_uint32), ("n", ctypes.c_uint32),
                ("cells_per_axis", ctypes.c_int32), ("cell_base", ctypes.c_uint32),
                ("field", ctypes.c_double), ("cell_edge", ctypes.c_double)]


class EngineError(RuntimeError):
    pass


def build() -> None:
    """Compile libboids_engine.so for gfx950 in the package directory."""
    subprocess.run(["make", "-s", "-f", os.path.join(PKG_DIR, "csrc", "Makefile")], check=True)


_lib = None

# every entry point of include/boids_engine.h
EXPORTS = ["bs_create", "bs_destroy", "bs_last_error", "bs_generate_init", "bs_read_init_file",
           "bs_set_state_f32", "bs_set_state_f64", "bs_step", "bs_sync", "bs_get_state_f32",
           "bs_get_state_f64", "bs_get_forces_f32", "bs_get_forces_f64", "bs_get_digest",
           "bs_get_info", "bs_get_timing", "bs_reset_timing", "bs_debug_thresholds",
           "bs_debug_rechecks", "bs_debug_copy", "bs_metrics", "bs_version", "bs_create_batch",
           "bs_get_flock_count", "bs_get_flock_info", "bs_metrics_flock", "bs_debug_grid",
           "bs_metrics_flock_noalloc"]


class GridInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("ncells", ctypes.c_uint32), ("nchunks", ctypes.c_uint32),
                ("max_chunks", ctypes.c_uint32), ("overflow", ctypes.c_uint32),
                ("x_columns_per_cell", ctypes.c_int32), ("chunk_boids", ctypes.c_int32),
                ("chunk_span", ctypes.c_int32)]


class Metrics(ctypes.Structure):
    _fields_ = [("polarisation", ctypes.c_double), ("mean_nn", ctypes.c_double),
                ("clusters", ctypes.c_uint64), ("largest", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load the engine library (raises if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P, c_int, vp = ctypes.POINTER, ctypes.c_int, ctypes.c_void_p
    f32p, f64p, u64p = P(ctypes.c_float), P(ctypes.c_double), P(ctypes.c_uint64)
    sig = {
        "bs_create": (c_int, [P(Params), P(vp)]),
        "bs_destroy": (None, [vp]),
        "bs_last_error": (ctypes.c_char_p, [vp]),
        "bs_generate_init": (c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, f32p]),
        "bs_read_init_file": (c_int, [ctypes.c_char_p, ctypes.c_uint32, f32p]),
        "bs_set_state_f32": (c_int, [vp, f32p]),
        "bs_set_state_f64": (c_int, [vp, f64p, f64p]),
        "bs_step": (c_int, [vp, ctypes.c_uint32]),
        "bs_sync": (c_int, [vp]),
        "bs_get_state_f32": (c_int, [vp, f32p, f32p]),
        "bs_get_state_f64": (c_int, [vp, f64p, f64p]),
        "bs_get_forces_f32": (c_int, [vp, f32p]),
        "bs_get_forces_f64": (c_int, [vp, f64p]),
        "bs_get_digest": (c_int, [vp, u64p]),
        "bs_get_info": (c_int, [vp, P(Info)]),
        "bs_get_timing": (c_int, [vp, ctypes.c_int32, f64p, u64p]),
        "bs_reset_timing": (c_int, [vp]),
        "bs_debug_thresholds": (c_int, [P(Params), f64p]),
        "bs_debug_rechecks": (c_int, [vp, u64p]),
        "bs_debug_copy": (c_int, [ctypes.c_uint64, ctypes.c_int32, f64p]),
        "bs_metrics": (c_int, [vp, ctypes.c_double, ctypes.c_uint32, P(Metrics)]),
        "bs_version": (ctypes.c_char_p, []),
        "bs_create_batch": (c_int, [P(Params), ctypes.c_uint32, P(vp)]),
        "bs_get_flock_count": (c_int, [vp, P(ctypes.c_uint32)]),
        "bs_get_flock_info": (c_int, [vp, ctypes.c_uint32, P(FlockInfo)]),
        "bs_metrics_flock": (c_int, [vp, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, P(Metrics)]),
        "bs_metrics_flock_noalloc": (c_int, [vp, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, P(Metrics)]),
        "bs_debug_grid": (c_int, [vp, P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32),
                                  P(ctypes.c_uint32), P(GridInfo)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def make_params(n, field, sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008),
                vmax=0.005, vmin=0.001, mode=MODE_GRID_F64, flags=0, cells_per_axis=0, device=0, block=0,
                tile=0, fill=0) -> Params:
    return Params(n, flags, field, *sep, *ali, *coh, vmax, vmin, mode, cells_per_axis, device, block, tile, fill)


def generate_init(kind: int, n: int, field: float, vmax: float, vmin: float, seed=None) -> np.ndarray:
    """N x 6 float32 initial state (the reference init-file layout).  seed None
    leaves drand48 unseeded (the all-pairs engine's behaviour)."""
    out = np.zeros((n, 6), np.float32)
    rc = lib().bs_generate_init(kind, 0 if seed is None else 1, 0 if seed is None else int(seed), n, field,
                                vmax, vmin, _p(out, ctypes.c_float))
    if rc != 0:
        raise EngineError(lib().bs_last_error(None).decode())
    return out


def debug_thresholds(p: Params) -> np.ndarray:
    out = np.zeros(6, np.float64)
    rc = lib().bs_debug_thresholds(ctypes.byref(p), _p(out, ctypes.c_double))
    if rc != 0:
        raise EngineError(lib().bs_last_error(None).decode())
    return out


def debug_copy(nbytes: int, reps: int = 1) -> float:
    """Runs `reps` plain device copies of `nbytes`; returns ms per copy."""
    ms = ctypes.c_double()
    rc = lib().bs_debug_copy(nbytes, reps, ctypes.byref(ms))
    if rc != 0:
        raise EngineError(lib().bs_last_error(None).decode())
    return ms.value


class Engine:
    """Owns one bs_engine handle (device buffers + one HIP stream)."""

    def __init__(self, params):
        """params: one Params (bs_create) or a sequence of Params, one per
        flock of a parameter sweep (bs_create_batch)."""
        h = ctypes.c_void_p()
        if isinstance(params, Params):
            self.params = params
            self.n = params.n
            rc = lib().bs_create(ctypes.byref(params), ctypes.byref(h))
            what = "bs_create"
        else:
            arr = (Params * len(params))(*params)
            self.params = list(params)
            self.n = sum(p.n for p in params)
            rc = lib().bs_create_batch(arr, len(params), ctypes.byref(h))
            what = "bs_create_batch"
        if rc != 0:
            raise EngineError(f"{what} failed ({rc}): {lib().bs_last_error(None).decode()}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            raise EngineError(f"{what} failed ({rc}): {lib().bs_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().bs_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, xyzv6=None, pos=None, vel=None):
        if xyzv6 is not None:
            a = np.ascontiguousarray(xyzv6, np.float32).reshape(self.n, 6)
            self._check(lib().bs_set_state_f32(self._h, _p(a, ctypes.c_float)), "bs_set_state_f32")
        else:
            p = np.ascontiguousarray(pos, np.float64).reshape(self.n, 3)
            v = np.ascontiguousarray(vel, np.float64).reshape(self.n, 3)
            self._check(lib().bs_set_state_f64(self._h, _p(p, ctypes.c_double), _p(v, ctypes.c_double)),
                        "bs_set_state_f64")

    def step(self, nsteps=1):
        self._check(lib().bs_step(self._h, nsteps), "bs_step")

    def sync(self):
        self._check(lib().bs_sync(self._h), "bs_sync")

    def state_f64(self):
        p = np.zeros((self.n, 3), np.float64)
        v = np.zeros((self.n, 3), np.float64)
        self._check(lib().bs_get_state_f64(self._h, _p(p, ctypes.c_double), _p(v, ctypes.c_double)),
                    "bs_get_state_f64")
        return p, v

    def state_f32(self):
        p = np.zeros((self.n, 3), np.float32)
        v = np.zeros((self.n, 3), np.float32)
        self._check(lib().bs_get_state_f32(self._h, _p(p, ctypes.c_float), _p(v, ctypes.c_float)),
                    "bs_get_state_f32")
        return p, v

    def forces_f64(self):
        f = np.zeros((self.n, 9), np.float64)
        self._check(lib().bs_get_forces_f64(self._h, _p(f, ctypes.c_double)), "bs_get_forces_f64")
        return f

    def forces_f32(self):
        f = np.zeros((self.n, 9), np.float32)
        self._check(lib().bs_get_forces_f32(self._h, _p(f, ctypes.c_float)), "bs_get_forces_f32")
        return f

    def digest(self):
        d = np.zeros((self.n, 4), np.uint64)
        self._check(lib().bs_get_digest(self._h, _p(d, ctypes.c_uint64)), "bs_get_digest")
        return d

    def info(self) -> Info:
        i = Info()
        self._check(lib().bs_get_info(self._h, ctypes.byref(i)), "bs_get_info")
        return i

    def timing(self, kernel):
        ms = ctypes.c_double()
        cnt = ctypes.c_uint64()
        self._check(lib().bs_get_timing(self._h, kernel, ctypes.byref(ms), ctypes.byref(cnt)), "bs_get_timing")
        return ms.value, cnt.value

    def reset_timing(self):
        self._check(lib().bs_reset_timing(self._h), "bs_reset_timing")

    def rechecks(self):
        out = ctypes.c_uint64()
        self._check(lib().bs_debug_rechecks(self._h, ctypes.byref(out)), "bs_debug_rechecks")
        return out.value

    def metrics(self, eps, min_size=2, flock=None) -> dict:
        """Polarisation, mean nearest-neighbour distance, cluster count (graph
        d < eps, components >= min_size) and largest cluster of the state
        (of flock `flock` of a batch)."""
        m = Metrics()
        if flock is None:
            self._check(lib().bs_metrics(self._h, eps, min_size, ctypes.byref(m)), "bs_metrics")
        else:
            self._check(lib().bs_metrics_flock(self._h, flock, eps, min_size, ctypes.byref(m)),
                        "bs_metrics_flock")
        return m.as_dict()

    def grid(self):
        """The grid of the last sort: (info, sorted cell keys, sorted ids, cell
        starts, chunk table as nchunks x 4)."""
        info = GridInfo()
        self._check(lib().bs_debug_grid(self._h, None, None, None, None, ctypes.byref(info)), "bs_debug_grid")
        u32 = ctypes.c_uint32
        key = np.zeros(info.n, np.uint32)
        ids = np.zeros(info.n, np.uint32)
        start = np.zeros(info.ncells + 1, np.uint32)
        table = np.zeros((max(info.max_chunks, 1), 4), np.uint32)
        self._check(lib().bs_debug_grid(self._h, _p(key, u32), _p(ids, u32), _p(start, u32), _p(table, u32),
                                        ctypes.byref(info)), "bs_debug_grid")
        return info, key, ids, start, table[:min(info.nchunks, info.max_chunks)].copy()

    def metrics_noalloc(self, eps, min_size=2, flock=0) -> dict:
        """metrics() of flock `flock` with no device allocation (the library's
        static scratch array; calls overlapping on one GPU race)."""
        m = Metrics()
        self._check(lib().bs_metrics_flock_noalloc(self._h, flock, eps, min_size, ctypes.byref(m)),
                    "bs_metrics_flock_noalloc")
        return m.as_dict()

    def flock_count(self) -> int:
        c = ctypes.c_uint32()
        self._check(lib().bs_get_flock_count(self._h, ctypes.byref(c)), "bs_get_flock_count")
        return c.value

    def flock_info(self, k) -> FlockInfo:
        i = FlockInfo()
        self._check(lib().bs_get_flock_info(self._h, k, ctypes.byref(i)), "bs_get_flock_info")
        return i


class BoidSimulation:
    """Mirror of the reference class (src/boid_simulation.h:15-71) over the C ABI.

    The variant is chosen by `mode` (the reference picks it at link time:
    boidsim = all-pairs, boidsim_grid = grid, CMakeLists.txt:85-103).
    Per-id getters read a host mirror refreshed once per update().
    """

    def __init__(self, mode=MODE_ALLPAIRS_F32REF, forces=False, device=0):
        self.mode = mode
        self._flags = FLAG_FORCES if forces else 0
        self.device = device
        self.N = 0
        self.time_step = 0
        self._engine = None
        self._mirror = None
        self._fmirror = None

    def setup(self, number_of_agents, field_size, separation_sight_distance, separation_sight_angle,
              separation_force_coefficient, alignment_sight_distance, alignment_sight_angle,
              alignment_force_coefficient, cohesion_sight_distance, cohesion_sight_angle,
              cohesion_force_coefficient, velocity_max, velocity_min, initialization, rand_seed):
        """src/boid_simulation.cpp:46-77 (velocity_max before velocity_min)."""
        self.N = number_of_agents
        self.field_size = self.field_size_X = self.field_size_Y = self.field_size_Z = field_size
        self.separation = (separation_sight_distance, separation_sight_angle, separation_force_coefficient)
        self.alignment = (alignment_sight_distance, alignment_sight_angle, alignment_force_coefficient)
        self.cohesion = (cohesion_sight_distance, cohesion_sight_angle, cohesion_force_coefficient)
        self.velocity = (velocity_max, velocity_min)
        self.initialization_type = initialization
        self.rand_seed = rand_seed
        self.time_step = 0
        p = make_params(number_of_agents, field_size, self.separation, self.alignment, self.cohesion,
                        velocity_max, velocity_min, mode=self.mode, flags=self._flags, device=self.device)
        self._engine = Engine(p)

    def init(self):
        """src/boid_simulation.cpp:79-137.  The all-pairs engine never seeds
        drand48; the grid engine calls srand48(seed) (src/boid_simulation_grid.cpp:98)."""
        seeded = self.mode in (MODE_GRID_F64, MODE_GRID_MIXED, MODE_GRID_F32REF)
        F = self.field_size
        vmax, vmin = self.velocity
        if self.initialization_type in ("test", "random_uniform"):
            kind = INIT_TEST if self.initialization_type == "test" else INIT_RANDOM_UNIFORM
            st = generate_init(kind, self.N, F, vmax, vmin, self.rand_seed if seeded else None)
        else:
            st = np.zeros((self.N, 6), np.float32)
            rc = lib().bs_read_init_file(self.initialization_type.encode(), self.N, _p(st, ctypes.c_float))
            if rc != 0:
                raise EngineError(lib().bs_last_error(None).decode())
        self._engine.set_state(xyzv6=st)
        self._refresh()

    def update(self):
        self._engine.step(1)
        self._refresh()

    def _refresh(self):
        self._mirror = self._engine.state_f32()
        self._fmirror = self._engine.forces_f32() if self._flags & FLAG_FORCES else None

    def get(self, id):
        return tuple(float(x) for x in self._mirror[0][id])

    def get_v(self, id):
        return tuple(float(x) for x in self._mirror[1][id])

    def get_force(self, id):
        """(float)w * dv per rule (src/boid_simulation.cpp:272-290)."""
        if self._fmirror is None:
            return (0.0,) * 9
        f = self._fmirror[id]
        w = [np.float32(self.cohesion[2])] * 3 + [np.float32(self.separation[2])] * 3 + \
            [np.float32(self.alignment[2])] * 3
        return tuple(float(np.float32(wk) * fk) for wk, fk in zip(w, f))

    def is_openmp_enabled(self):
        return False

    def get_max_threads(self):
        return -1
