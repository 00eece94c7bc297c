"""Ensembles: independent porous boxes of different sizes, gases and seeds
advanced by one handle (one set of launches, one captured graph per step
pair).  A member of an ensemble must step exactly as it does alone."""
import numpy as np
import pytest

from master_amd import config as C

SPECS = [
    dict(edge=32, porosity=0.5, radius=4.0, kn=0.1, mach=0.05, seed=11, per_voxel=8.0),
    dict(edge=48, porosity=0.7, radius=3.0, kn=0.03, mach=0.02, seed=12, per_voxel=6.0),
    dict(edge=32, porosity=0.4, radius=2.5, kn=0.5, mach=0.08, seed=13, per_voxel=8.0),
]


def _member(dsmc, spec):
    e, r = spec["edge"], spec["radius"]
    w = dsmc.WorldData.poly_spheres((e, e, e), r, r, spec["porosity"], spec["seed"])
    return C.member_settings(spec, w.porosity), w


def test_ensemble_specs_are_deterministic_and_in_range():
    a, b = C.ensemble_members(), C.ensemble_members()
    assert a == b and len(a) == 64
    assert [m["seed"] for m in a] == list(range(100, 164))
    for m in a:
        assert m["edge"] in C.ENSEMBLE_EDGES
        assert 0.3 <= m["porosity"] <= 0.8 and 2.0 <= m["radius"] <= 8.0
        assert 0.01 <= m["kn"] <= 1.0 and 0.01 <= m["mach"] <= 0.1
    assert len({m["edge"] for m in a}) == len(C.ENSEMBLE_EDGES)


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_member_equals_solo_run(gpu, graph):
    dsmc = gpu
    members = [_member(dsmc, s) for s in SPECS]
    ens = dsmc.System([m[0] for m in members], [m[1] for m in members])
    n_ens = ens.initialize()
    ens.set_sampling(1)
    ens.set_option(dsmc.OPT_USE_GRAPH, int(graph))
    ens.step(20)
    st = ens.state()
    ce = ens.counters()
    first = 0
    coll = wall = 0
    for ini, w in members:
        g = dsmc.System(ini, w)
        n = g.initialize()
        g.set_sampling(1)
        g.set_option(dsmc.OPT_USE_GRAPH, int(graph))
        g.step(20)
        solo = g.state()
        c = g.counters()
        coll += c["collisions"]
        wall += c["wall_collisions"]
        assert np.array_equal(st[first:first + n], solo), f"member at {first} differs"
        first += n
        g.close()
    assert first == n_ens
    assert ce["collisions"] == coll and ce["wall_collisions"] == wall
    assert ce["errors"] == 0


@pytest.mark.gpu
def test_tuning_knobs_leave_results_unchanged(gpu):
    """Collide chunking, graph length, the fused gather + collide pass, the
    mover's wall-path queue (0: every wall-path particle goes through the
    overflow pass), the LDS capacity (1024: the fused pass's large-batch
    variant; 96: most cells on the fused pass's global-memory path, with the
    move fused in or not), moving ahead (the fused pass moving for the next step, on by
    default; eager chains and graphs) and the autotuner's choice change
    launch shapes only: trajectories stay bit-identical."""
    dsmc = gpu
    members = [_member(dsmc, s) for s in SPECS[:2]]

    def make():
        g = dsmc.System([m[0] for m in members], [m[1] for m in members])
        g.initialize()
        g.set_sampling(1)
        g.set_option(dsmc.OPT_USE_GRAPH, 1)
        return g

    tuned = make()
    rep = tuned.autotune(4)
    assert rep["trials"] >= 6 and rep["ms_per_step"] > 0
    assert rep["ms_per_step"] <= rep["ms_per_step_default"]
    tuned.step(10)
    total = tuned.counters()["steps"]
    ref = (tuned.state(), tuned.moments()[0], tuned.counters()["collisions"], tuned.counters()["wall_collisions"])
    tuned.close()
    for knobs in ({}, {dsmc.OPT_COLLIDE_CHUNK: 1, dsmc.OPT_GRAPH_STEPS: 4},
                  {dsmc.OPT_COLLIDE_CHUNK: 16, dsmc.OPT_GRAPH_STEPS: 8},
                  {dsmc.OPT_FUSED: 0}, {dsmc.OPT_FUSED: 0, dsmc.OPT_USE_GRAPH: 0},
                  {dsmc.OPT_MOVE_QUEUE: 0}, {dsmc.OPT_MOVE_QUEUE: 3, dsmc.OPT_USE_GRAPH: 0},
                  {dsmc.OPT_LDS_CAPACITY: 1024}, {dsmc.OPT_LDS_CAPACITY: 96},
                  {dsmc.OPT_LDS_CAPACITY: 96, dsmc.OPT_USE_GRAPH: 0}, {dsmc.OPT_MOVE_AHEAD: 0},
                  {dsmc.OPT_MOVE_AHEAD: 0, dsmc.OPT_USE_GRAPH: 0}, {dsmc.OPT_USE_GRAPH: 0},
                  {dsmc.OPT_GRAPH_STEPS: 16}):
        g = make()
        for k, v in knobs.items():
            g.set_option(k, v)
        g.step(total)
        assert g.counters()["steps"] == total
        assert np.array_equal(g.state(), ref[0]), knobs
        assert np.array_equal(g.moments()[0], ref[1]), knobs
        assert g.counters()["collisions"] == ref[2]
        assert g.counters()["wall_collisions"] == ref[3]
        g.close()
