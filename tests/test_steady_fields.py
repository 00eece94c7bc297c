"""Steady-state per-cell fields against the reference solver's ensemble.

The fixture (tests/golden/reference_steady.npz, tests/golden/make_steady.py)
holds, for 4 seeds of the compiled reference on the CPU-size workload (32^3
packed spheres r=4, 8^3 cells, ~200k particles, Kn 0.1, diffuse walls, body
force along x), per-cell sums of (1, v, v^2) over 12 windows of 500 steps.
The GPU runs the same configuration for other seeds; after 1000 steps of
equilibration, per-cell density, drift velocity and temperature must agree
within the ensemble standard error (z-test in tests/steady_stats.py).
"""
import os

import numpy as np
import pytest

from master_amd import config as C
from tests import dsmc_cases as K
from tests import steady_stats as S

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_steady.npz")
K0 = 2                      # windows skipped (equilibration)


def _reference():
    d = np.load(FIXTURE, allow_pickle=False)
    seeds = [int(s) for s in d["seeds"]]
    return d, seeds, [d[f"sums_{s}"] for s in seeds]


def _mass():
    w = K.world()
    ini = K.settings(w, 200_000, seed=1)
    _, m, _ = C.gas_scales(ini)
    return m


def test_steady_method_is_calibrated_and_sensitive():
    """Reference seeds split 2 + 2 agree; a planted 1% temperature bias or a
    drift of 2e-3 (reduced units) in one half is detected."""
    d, seeds, runs = _reference()
    W, m = int(d["window"]), _mass()
    for a, b in (([0, 1], [2, 3]), ([0, 3], [1, 2])):
        FA = S.ensemble([runs[i] for i in a], m, K0, W)
        FB = S.ensemble([runs[i] for i in b], m, K0, W)
        z, act = S.z_scores(FA, FB)
        ok, summary = S.consistent(z)
        assert ok, summary
        assert act.sum() > 400
        hot = FB.copy()
        hot[..., 4] *= 1.01
        ok, summary = S.consistent(S.z_scores(FA, hot)[0])
        assert not ok and summary["T"]["mean_z2"] > 1.6, summary["T"]
        drift = FB.copy()
        drift[..., 1] += 2e-3
        ok, summary = S.consistent(S.z_scores(FA, drift)[0])
        assert not ok and summary["ux"]["mean_z2"] > 1.6, summary["ux"]


@pytest.mark.gpu
def test_steady_fields_match_reference(gpu):
    d, seeds, runs = _reference()
    W, steps, m = int(d["window"]), int(d["steps"]), _mass()
    nwin = steps // W
    w = K.world()
    gpu_runs, coll, wall = [], [], []
    for seed in range(101, 109):
        ini = K.settings(w, 200_000, seed=seed)
        g = K.gpu_system(ini, w)
        g.initialize()
        g.set_sampling(1)
        g.set_option(gpu.OPT_USE_GRAPH, 1)
        sums = np.zeros((nwin, g.ncells, 7))
        for k in range(nwin):
            g.reset_moments()
            g.step(W)
            mom, ns = g.moments()
            assert ns == W
            sums[k] = mom
        c = g.counters()
        assert c["errors"] == 0 and c["steps"] == steps
        gpu_runs.append(sums)
        coll.append(c["collisions"])
        wall.append(c["wall_collisions"])
        g.close()
        print(f"seed {seed}: collisions {coll[-1]} wall hits {wall[-1]}", flush=True)
    FG = S.ensemble(gpu_runs, m, K0, W)
    FR = S.ensemble(runs, m, K0, W)
    z, act = S.z_scores(FG, FR)
    ok, summary = S.consistent(z)
    print(summary)
    assert ok, summary
    assert act.sum() > 400
    # whole-run totals: collisions and wall hits per run, GPU seeds vs reference seeds
    for name, g_vals in (("collisions", coll), ("wall", wall)):
        r_vals = np.array([int(d[f"{name}_{s}"]) for s in seeds], dtype=float)
        g_vals = np.array(g_vals, dtype=float)
        v = (g_vals.var(ddof=1) * (len(g_vals) - 1) + r_vals.var(ddof=1) * (len(r_vals) - 1)) / \
            (len(g_vals) + len(r_vals) - 2)
        zt = (g_vals.mean() - r_vals.mean()) / np.sqrt(v * (1 / len(g_vals) + 1 / len(r_vals)))
        assert abs(zt) < 4.0, (name, g_vals.mean(), r_vals.mean(), zt)
