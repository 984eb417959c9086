"""Exactness of the GPU grid build (the counting sort and the chunk table that
replace the reference's per-step cell-list registration,
src/boid_simulation_grid.cpp:383-451) against a CPU restatement:

  * cell keys from the FP64 state with the engine's binning (x at
    x_columns_per_cell columns per cell edge, y and z at the cell edge);
  * sorted order = a stable counting sort by (cell key, boid id);
  * cell starts = the exclusive prefix sum of the cell counts;
  * chunk table = every (y,z) row cut into span groups of at most 62 x
    columns, each group into runs of at most chunk_boids sorted boids, cut
    short where a run would cover more than chunk_span columns;
  * the table never exceeds its capacity (no overflow reported).
"""
import sys

import numpy as np
import pytest

import boidssimulatormpi_amd as B
from boidssimulatormpi_amd import engine as E

from grid_model import expected_grid

pytestmark = pytest.mark.gpu

CFG0 = dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(e, flocks, label):
    info, key, ids, start, table = e.grid()
    assert info.overflow == 0 and info.nchunks <= info.max_chunks, (label, info.nchunks, info.max_chunks)
    pos, _ = e.state_f64()
    ek, eids, estart, etable = expected_grid(pos, flocks, info.x_columns_per_cell, info.chunk_boids,
                                             info.chunk_span)
    assert np.array_equal(key, ek), label
    assert np.array_equal(ids, eids), label
    assert np.array_equal(start, estart), label
    assert info.nchunks == etable.shape[0] and np.array_equal(table, etable), (label, info.nchunks, etable.shape)
    occ = np.diff(estart)
    say(f"{label}: {info.n} boids, {info.ncells} cells (max occupancy {occ.max()}), {info.nchunks} chunks "
        f"of capacity {info.max_chunks}, chunk <= {info.chunk_boids} boids / {info.chunk_span} columns")


def single(e, n, F):
    i = e.info()
    return [(0, n, i.cells_per_axis, F, 0, 0)]


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_grid_build_uniform_and_evolved(mode):
    n, F = 1 << 20, 4.0
    st = B.generate_init(E.INIT_ISOTROPIC, n, F, CFG0["vmax"], CFG0["vmin"], seed=2)
    e = B.Engine(B.make_params(n, F, **CFG0, mode=mode))
    e.set_state(xyzv6=st)
    check(e, single(e, n, F), f"{E.MODE_NAMES[mode]} t=0")       # the sort of the uploaded state
    e.step(1)
    check(e, single(e, n, F), f"{E.MODE_NAMES[mode]} t=1")       # keys from the fused kernel's epilogue
    e.step(199)
    check(e, single(e, n, F), f"{E.MODE_NAMES[mode]} t=200")     # clustered by cohesion
    e.close()


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_grid_build_clustered_blobs(mode):
    # 64 Gaussian blobs at the 2^24-boid configuration's density: cells of
    # very uneven occupancy, rows with long sparse stretches (short cuts)
    n = 1 << 20
    F = (n / (2.0 ** 24 / 1000.0)) ** (1.0 / 3.0)
    st = B.generate_init(E.INIT_CLUSTERED, n, F, CFG0["vmax"], CFG0["vmin"], seed=4)
    e = B.Engine(B.make_params(n, F, **CFG0, mode=mode))
    e.set_state(xyzv6=st)
    check(e, single(e, n, F), f"blobs {E.MODE_NAMES[mode]} t=0")
    e.step(3)
    check(e, single(e, n, F), f"blobs {E.MODE_NAMES[mode]} t=3")
    e.close()


def test_grid_build_batch_and_small_flocks():
    # a flock table: flock bases of keys, rows and boids; small flocks run
    # with smaller chunks (automatic fill)
    from boidssimulatormpi_amd import sweep as S
    jobs = S.generated_sweep(24, 11)
    e = B.Engine([j.params(E.MODE_GRID_MIXED, 0) for j in jobs])
    e.set_state(xyzv6=np.concatenate([S.initial_state(j) for j in jobs]))
    flocks, rb = [], 0
    for k in range(e.flock_count()):
        fi = e.flock_info(k)
        flocks.append((fi.boid_base, fi.n, fi.cells_per_axis, fi.field, fi.cell_base, rb))
        rb += fi.cells_per_axis ** 2
    check(e, flocks, "batch of 24 t=0")
    e.step(5)
    check(e, flocks, "batch of 24 t=5")
    e.close()


def test_cells_per_axis_is_the_exact_cube_root_cap():
    # the cells-per-boid cap of the automatic grid is the exact integer cube
    # root (perfect cubes included), as the autotuner assumes
    from boidssimulatormpi_amd import autotune as T
    for n in (64, 4096, 32768, 1 << 18, 35937):
        p = B.make_params(n, 20.0, **CFG0, mode=E.MODE_GRID_F64)
        e = B.Engine(p)
        assert e.info().cells_per_axis == T.auto_cells(p) == T.icbrt(n), n
        e.close()


@pytest.mark.parametrize("mode", [E.MODE_GRID_F64, E.MODE_GRID_MIXED])
def test_grid_build_crowded_cells_span_sort_windows(mode):
    # cells of 3000, 1500, 700, 100, 65 and 64 boids (ids drawn at random):
    # a cell of more than 64 boids takes its in-cell id order from 1024-slot
    # sorted windows merged by binary search, a smaller one from a pass over
    # its ids; a cell spanning three or four windows must still come out in
    # ascending id order
    n, F = 1 << 16, 1.0
    rng = np.random.default_rng(11)
    st = B.generate_init(E.INIT_RANDOM_UNIFORM, n, F, CFG0["vmax"], CFG0["vmin"], seed=5)
    e = B.Engine(B.make_params(n, F, **CFG0, mode=mode))
    G = e.info().cells_per_axis
    edge = F / G
    ids = rng.permutation(n)
    at = 0
    for size, cell in ((3000, (3, 4, 5)), (1500, (7, 7, 7)), (700, (0, G - 1, 2)), (100, (9, 1, 1)),
                       (65, (5, 9, 3)), (64, (1, 5, 9))):
        sel = ids[at:at + size]
        at += size
        lo = (np.array(cell) + 0.3) * edge
        st[sel, 0:3] = lo + rng.random((size, 3)) * 0.2 * edge
    e.set_state(xyzv6=st)
    check(e, single(e, n, F), f"crowded {E.MODE_NAMES[mode]} t=0")
    e.step(2)
    check(e, single(e, n, F), f"crowded {E.MODE_NAMES[mode]} t=2")
    e.close()
