"""CPU restatement of the engine's grid build, the checker of
tests/test_gpu_grid_build.py (and pinned by tests/test_grid_model.py):
cell keys from the FP64 state, a stable counting sort by (cell key, boid
id), cell starts, and the neighbour kernels' chunk table.  The reference's
counterpart is the per-step cell-list registration of
src/boid_simulation_grid.cpp:383-451."""
import numpy as np


def axis_cell(x, inv, G):
    c = np.floor(x * inv)
    c = np.where(c >= 0.0, c, 0.0)          # also NaN
    return np.minimum(c, G - 1).astype(np.int64)


def expected_grid(pos, flocks, sx, chunk, cspan):
    """flocks: list of (boid_base, n, G, field, cell_base, row_base)."""
    n = pos.shape[0]
    key = np.zeros(n, np.int64)
    for (b0, m, G, F, cb, _) in flocks:
        p = pos[b0:b0 + m]
        inv = G / F
        Gx = sx * G
        cx = axis_cell(p[:, 0], sx * inv, Gx)
        cy = axis_cell(p[:, 1], inv, G)
        cz = axis_cell(p[:, 2], inv, G)
        key[b0:b0 + m] = cb + cx + Gx * (cy + G * cz)
    ncells = sum(sx * G ** 3 for (_, _, G, _, _, _) in flocks)
    order = np.lexsort((np.arange(n), key))           # stable: by key, then id
    counts = np.bincount(key, minlength=ncells)
    start = np.zeros(ncells + 1, np.int64)
    start[1:] = np.cumsum(counts)
    table = []
    for f, (_, _, G, _, cb, _) in enumerate(flocks):
        Gx = sx * G
        ng = -(-Gx // 62)
        span = -(-Gx // ng)
        for rl in range(G * G):
            cs = start[cb + rl * Gx: cb + rl * Gx + Gx + 1]
            for x0 in range(0, Gx, span):
                x1 = min(Gx, x0 + span)
                hi = cs[x1]
                t0, xa = cs[x0], x0
                while t0 < hi:
                    while cs[xa + 1] <= t0:
                        xa += 1
                    t1 = min(hi, t0 + chunk)
                    if xa + cspan < x1:
                        t1 = min(t1, cs[xa + cspan])
                    table.append((rl, t0, t1, f))
                    t0 = t1
    return key[order], order, start, np.array(table, np.int64).reshape(-1, 4)
