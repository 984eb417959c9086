"""Long-run statistics of both GPU grid paths pinned to the reference's own
dynamics (the north star's bar for chaotic runs: polarisation, mean
nearest-neighbour distance and cluster count within 1 %).

The fixture tests/golden/longrun_n16384.npz (written by
tests/golden/make_longrun.py) holds the time series of those statistics for
the reference's float dynamics -- the float oracle, bit-identical to the
reference's grid engine (checked there against oracle/_ref/ref_grid over the
whole run) -- at N = 16384, F = 1, T = 2000, from the reference's
random_uniform init at srand48(seed) for six seeds, with two rule sets: the
configured rules (disordered flock) and an ordered regime (strong alignment,
polarisation ~0.9).  The GPU runs the same initial states in FP64 plain and
in mixed precision; pointwise the runs diverge (chaos), so the comparison is
between time averages over the second half of the runs and the six seeds.

At t = 0 the states are identical, so there the GPU metrics must equal the
fixture's exactly (same init, same metric definitions).
"""
import os
import sys

import numpy as np
import pytest

import boidssimulatormpi_amd as B
from boidssimulatormpi_amd import engine as E

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "longrun_n16384.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rules_of(a):
    sr, sa, sw, ar, aa, aw, cr, ca, cw, vmax, vmin = [float(x) for x in a]
    return dict(sep=(sr, sa, sw), ali=(ar, aa, aw), coh=(cr, ca, cw), vmax=vmax, vmin=vmin)


def gpu_series(mode, rules, n, F, seed, T, every, eps, min_size):
    e = B.Engine(B.make_params(n, F, **rules, mode=mode))
    e.set_state(xyzv6=B.generate_init(E.INIT_RANDOM_UNIFORM, n, F, rules["vmax"], rules["vmin"], seed=seed))
    out = []
    for t in range(0, T + 1, every):
        if t:
            e.step(every)
        m = e.metrics(eps, min_size)
        out.append([t, m["polarisation"], m["mean_nn"], m["clusters"], m["largest"]])
    e.close()
    return np.array(out)


@pytest.mark.parametrize("ruleset", ["cfg", "ordered"])
def test_long_run_statistics_match_reference(ruleset):
    if not os.path.exists(FIXTURE):
        pytest.skip("fixture tests/golden/longrun_n16384.npz not generated")
    g = np.load(FIXTURE)
    n, F, T, every = int(g["n"]), float(g["field"]), int(g["steps"]), int(g["every"])
    eps, min_size = float(g["eps"]), int(g["min_size"])
    rules = rules_of(g[f"{ruleset}_rules"])
    ref = g[f"{ruleset}_series"]                       # seeds x samples x (t, pol, nn, clusters, largest)
    late = ref[0, :, 0] >= T // 2
    want = ref[:, late, 1:4].mean((0, 1))
    for mode in (E.MODE_GRID_F64, E.MODE_GRID_MIXED):
        runs = []
        for k, seed in enumerate(g["seeds"]):
            s = gpu_series(mode, rules, n, F, int(seed), T, every, eps, min_size)
            assert np.array_equal(s[:, 0], ref[k, :, 0])
            # t = 0: the same state, so the same statistics
            assert s[0, 3] == ref[k, 0, 3] and s[0, 4] == ref[k, 0, 4], (seed, s[0], ref[k, 0])
            assert abs(s[0, 1] - ref[k, 0, 1]) < 1e-12 and abs(s[0, 2] - ref[k, 0, 2]) < 1e-12 * F
            runs.append(s)
        got = np.stack(runs)[:, late, 1:4].mean((0, 1))
        rel = np.abs(got - want) / np.abs(want)
        say(f"{ruleset} {E.MODE_NAMES[mode]}: polarisation {got[0]:.5f} (reference {want[0]:.5f}), mean NN "
            f"{got[1]:.6f} ({want[1]:.6f}), clusters {got[2]:.1f} ({want[2]:.1f}); relative differences "
            f"{rel[0]:.4f} {rel[1]:.4f} {rel[2]:.4f}")
        assert rel[1] <= 0.01 and rel[2] <= 0.01, (ruleset, E.MODE_NAMES[mode], got, want)
        if ruleset == "ordered":
            assert want[0] > 0.8 and rel[0] <= 0.01, (got[0], want[0])
        else:
            # disordered: polarisation is the random-direction value ~1/sqrt(N) in both
            assert got[0] < 3 / np.sqrt(n) and want[0] < 3 / np.sqrt(n), (got[0], want[0])
