"""The allocation-free entry point bs_metrics_flock_noalloc (static device
scratch) through the C ABI: on two calls in a row, the second on a smaller
flock, it returns exactly what bs_metrics_flock (temporary workspace)
returns, and it rejects a flock whose workspace exceeds the static array."""
import ctypes
import sys

import numpy as np
import pytest

import boidssimulatormpi_amd as B
from boidssimulatormpi_amd import engine as E

pytestmark = pytest.mark.gpu

CFG0 = dict(sep=(0.02, 0.75, 0.002), ali=(0.05, 0.75, 0.06), coh=(0.08, 0.75, 0.008), vmax=0.005, vmin=0.001)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_noalloc_metrics_equal_allocating_ones():
    runs = [(65536, (65536 / 16384.0) ** (1 / 3), E.MODE_GRID_MIXED, 7),
            (6000, 0.72, E.MODE_GRID_F64, 3)]          # the second call: a smaller flock
    for n, F, mode, seed in runs:
        e = B.Engine(B.make_params(n, F, **CFG0, mode=mode))
        e.set_state(xyzv6=B.generate_init(E.INIT_ISOTROPIC, n, F, CFG0["vmax"], CFG0["vmin"], seed=seed))
        e.step(30)
        for eps, ms in ((0.02, 2), (0.05, 1)):
            a = e.metrics(eps, ms)
            b = e.metrics_noalloc(eps, ms)
            say(f"n={n} eps={eps}: {b}")
            assert a == b, (n, eps, a, b)
        e.close()
    # a batch: flock by flock, the largest first
    from boidssimulatormpi_amd import sweep as S
    jobs = S.generated_sweep(6, 5)
    sw = S.Sweep(jobs, mode=E.MODE_GRID_MIXED, flags=0)
    sw.set_initial_state([S.initial_state(j) for j in jobs])
    sw.step(5)
    for k in sorted(range(len(jobs)), key=lambda k: -jobs[k].n):
        assert sw.engine.metrics(0.02, 2, flock=k) == sw.engine.metrics_noalloc(0.02, 2, flock=k), k
    sw.close()


def test_noalloc_rejects_a_flock_beyond_the_scratch():
    n, F = 1 << 20, 4.0
    e = B.Engine(B.make_params(n, F, **CFG0, mode=E.MODE_GRID_MIXED))
    e.set_state(xyzv6=B.generate_init(E.INIT_ISOTROPIC, n, F, CFG0["vmax"], CFG0["vmin"], seed=2))
    m = E.Metrics()
    rc = E.lib().bs_metrics_flock_noalloc(e._h, 0, 0.02, 2, ctypes.byref(m))
    assert rc == -1 and b"static scratch" in E.lib().bs_last_error(e._h)
    e.close()
