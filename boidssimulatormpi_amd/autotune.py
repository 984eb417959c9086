"""Autotuner for the grid modes: picks the neighbour-kernel workgroup size,
the LDS tile depth, the chunk fill and the cell size for a workload by
timing a few updates of each candidate on the workload's own state.

The reference has no counterpart (its grid size is fixed by GRID_NUM,
src/boid_simulation_grid.cpp:80-95); here every choice keeps the results
within the parity bars (FP64: neighbour sets exact; the summation order,
and so the last bits, depend on the cell size), so the tuner only trades
speed.

Results are cached in a JSON file (default
~/.cache/boidssimulatormpi_amd/tuning.json, override with
BOIDS_TUNING_CACHE) keyed by the workload's shape: mode, flock count,
total boids, and a digest of every flock's (n, field, radii).
"""
from __future__ import annotations

import hashlib
import itertools
import json
import math
import os
import time

from . import engine as E

CACHE_PATH = os.environ.get("BOIDS_TUNING_CACHE") or os.path.join(
    os.path.expanduser("~"), ".cache", "boidssimulatormpi_amd", "tuning.json")

BLOCKS = (64, 128, 256)
TILES = (5, 6, 9, 12)       # LDS candidates staged per thread and pass
FILLS = (60, 80, 100)       # chunk fill target, percent of the workgroup
CELL_SCALES = (1.0, 0.85)   # cells per axis relative to the automatic choice


def auto_cells(p: E.Params) -> int:
    """The engine's automatic cells per axis (boids_engine.hip: cells_per_axis)."""
    rmax = max(p.coh_r, p.ali_r, p.sep_r)
    if not (rmax > 0 and p.field > 0):
        return 0
    g = math.floor(p.field / (rmax * 1.0001))
    if g < 3:
        return 0
    cap = max(3, icbrt(max(p.n, 27)))
    return int(min(g, cap, 1024))


def icbrt(n: int) -> int:
    """Largest g with g**3 <= n (exact for perfect cubes)."""
    g = int(round(n ** (1.0 / 3.0)))
    while g ** 3 > n:
        g -= 1
    while (g + 1) ** 3 <= n:
        g += 1
    return g


def _with(p: E.Params, **kw) -> E.Params:
    q = E.Params.from_buffer_copy(p)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def workload_key(params) -> str:
    ps = [params] if isinstance(params, E.Params) else list(params)
    h = hashlib.sha256()
    for p in ps:
        h.update(repr((p.n, round(p.field, 12), round(p.sep_r, 12), round(p.ali_r, 12), round(p.coh_r, 12),
                       p.cells_per_axis)).encode())
    return f"mode{ps[0].mode}_flocks{len(ps)}_n{sum(p.n for p in ps)}_{h.hexdigest()[:16]}"


def candidates(params, blocks=BLOCKS, tiles=TILES, fills=FILLS, cell_scales=CELL_SCALES):
    """Every (block, tile, fill, cell scale) combination, as parameter lists."""
    ps = [params] if isinstance(params, E.Params) else list(params)
    for b, t, f, s in itertools.product(blocks, tiles, fills, cell_scales):
        out = []
        for p in ps:
            cells = p.cells_per_axis
            if s != 1.0:
                g0 = auto_cells(p) if cells == 0 else cells
                g = int(math.floor(g0 * s))
                if g < 3 or g == g0:
                    out = None
                    break
                cells = g
            out.append(_with(p, block=b, tile=t, fill=f, cells_per_axis=cells))
        if out is not None:
            yield {"block": b, "tile": t, "fill": f, "cell_scale": s}, out


def _time(ps, state, steps, warmup):
    eng = E.Engine(ps[0] if len(ps) == 1 else ps)
    try:
        eng.set_state(**state)
        eng.step(warmup)
        eng.sync()
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        return (time.perf_counter() - t0) / steps
    finally:
        eng.close()


def tune(params, state: dict, steps=8, warmup=2, cache=True, log=None, **space):
    """Returns (best tuning dict, best parameter list, all timings).

    params: one Params or a list (one per flock); state: keyword arguments of
    Engine.set_state (xyzv6=... or pos=..., vel=...) for the whole engine."""
    key = workload_key(params)
    db = {}
    if cache and os.path.exists(CACHE_PATH):
        try:
            with open(CACHE_PATH) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
    cands = list(candidates(params, **space))
    if cache and key in db:
        best = db[key]["best"]
        for t, ps in cands:
            if t == best:
                return t, ps, db[key]["timings"]
    timings = []
    for t, ps in cands:
        sec = _time(ps, state, steps, warmup)
        timings.append({**t, "ms_per_step": 1e3 * sec})
        if log:
            log(f"[tune] {t}: {1e3 * sec:.3f} ms/step")
    best_t = min(timings, key=lambda r: r["ms_per_step"])
    best = {k: best_t[k] for k in ("block", "tile", "fill", "cell_scale")}
    if cache:
        db[key] = {"best": best, "timings": timings}
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        tmp = CACHE_PATH + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1)
        os.replace(tmp, CACHE_PATH)
    for t, ps in cands:
        if t == best:
            return t, ps, timings
    raise RuntimeError("no candidate")
