"""CPU checks of the grid-build restatement (tests/grid_model.py) that the
GPU build is compared with: the counting sort against a brute-force sort,
and the chunk table's invariants (every sorted boid in exactly one chunk,
chunks inside one row, at most chunk_boids boids and chunk_span columns)."""
import numpy as np
import pytest

from grid_model import axis_cell, expected_grid


def test_axis_cell_clamps_and_rejects_nan():
    x = np.array([-1e-300, 0.0, 0.999, 1.0, 3.99999, 4.0, np.nan])
    assert list(axis_cell(x, 49 / 4.0, 49)) == [0, 0, 12, 12, 48, 48, 0]


@pytest.mark.parametrize("sx,chunk,cspan,G,n", [(2, 64, 12, 9, 4000), (1, 16, 16, 5, 300), (2, 48, 3, 4, 2000)])
def test_expected_grid_invariants(sx, chunk, cspan, G, n):
    rng = np.random.default_rng(G * 1000 + n)
    F = 1.0
    pos = rng.random((n, 3)) * F
    pos[: n // 4] = 0.3 + 0.02 * rng.standard_normal((n // 4, 3))    # a dense blob
    pos %= F
    key, ids, start, table = expected_grid(pos, [(0, n, G, F, 0, 0)], sx, chunk, cspan)
    # the sort: keys ascending, ids ascending within a key, a permutation
    assert np.all(np.diff(key) >= 0) and sorted(ids) == list(range(n))
    same = key[1:] == key[:-1]
    assert np.all(ids[1:][same] > ids[:-1][same])
    brute = sorted(range(n), key=lambda i: (int(np.floor(pos[i, 0] * sx * G / F)) + sx * G * (
        int(np.floor(pos[i, 1] * G / F)) + G * int(np.floor(pos[i, 2] * G / F))), i))
    assert list(ids) == brute
    assert start[0] == 0 and start[-1] == n and np.all(np.diff(start) >= 0)
    assert np.array_equal(np.diff(start), np.bincount(key, minlength=sx * G ** 3))
    # the chunks tile [0, n) in order, each inside one row, within the caps
    Gx = sx * G
    assert table[0, 1] == 0 and table[-1, 2] == n
    assert np.all(table[1:, 1] == table[:-1, 2]) and np.all(table[:, 2] > table[:, 1])
    assert np.all(table[:, 2] - table[:, 1] <= chunk)
    for rl, t0, t1, f in table:
        cols = key[t0:t1] % Gx
        assert f == 0 and np.all(key[t0:t1] // Gx == rl)
        assert cols.max() - cols.min() < cspan
