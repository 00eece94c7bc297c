"""Edge cases of the step on the GPU: no particles, one particle, and a cell
far above the collide pass's LDS capacity (4096 particles at most), which
takes the global-memory path of every collide variant.  With specular walls
and no body force the dense cell's collisions and wall reflections keep the
kinetic energy to FP32 round-off, and every launch shape gives the same
trajectory bit for bit."""
import numpy as np
import pytest

from oracle import movie_oracle as M
from tests import dsmc_cases as K


@pytest.mark.gpu
def test_empty_system_steps(gpu, tmp_path):
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000)
    g = dsmc.System(ini, w)
    g.set_movie(tmp_path / "movie.bin", 2, 8)
    g.load_state(np.zeros((0, 6)))
    g.step(5)
    c = g.counters()
    assert c["steps"] == 5 and c["collisions"] == 0 and c["errors"] == 0
    assert g.state().shape == (0, 6)
    g.set_movie(None)
    frames = M.read_frames(open(tmp_path / "movie.bin", "rb").read())
    assert len(frames) == 3 and not any(f.any() for f in frames)
    g.close()


@pytest.mark.gpu
def test_single_particle_flies_freely(gpu):
    """One particle in an empty voxel, no body force: until it meets a wall
    it moves v dt per step, and it never collides."""
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000, gravity=0.0)
    n = w["n"][0]
    L, dt = float(ini["Lx"]), float(ini["dt"])
    h = L / n
    empty = np.argwhere(w["types"].reshape(n, n, n) == 0)
    ijk = empty[len(empty) // 2]
    v = np.array([0.05, -0.03, 0.02]) * h / dt       # a few hundredths of a voxel per step
    rv = np.concatenate([(ijk + 0.5) * h, v])[None, :]
    g = dsmc.System(ini, w)
    g.load_state(rv)
    for _ in range(3):
        g.step(1)
        s = g.state()
        assert s.shape == (1, 6)
        if g.counters()["wall_collisions"] == 0:
            d = (s[0, :3] - rv[0, :3] + L / 2) % L - L / 2
            steps = g.counters()["steps"]
            assert np.allclose(d, v * dt * steps, rtol=1e-5, atol=1e-6 * h)
    assert g.counters()["collisions"] == 0
    g.close()


def _dense_cell_state(w, ini, count, seed=3):
    """`count` particles in one empty voxel, Maxwellian at the gas temperature."""
    n = w["n"][0]
    L = float(ini["Lx"])
    h = L / n
    empty = np.argwhere(w["types"].reshape(n, n, n) == 0)
    rng = np.random.default_rng(seed)
    ijk = empty[rng.integers(len(empty))]
    r = (ijk + rng.random((count, 3))) * h
    T, m = K.C.gas_scales(ini)[:2]
    v = rng.normal(0.0, np.sqrt(T / m), (count, 3))
    return np.concatenate([r, v], axis=1)


@pytest.mark.gpu
def test_dense_cell_takes_the_global_path(gpu):
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000, wall="specular", gravity=0.0)
    rv = _dense_cell_state(w, ini, 5000)
    e0 = np.sum(rv[:, 3:] ** 2)
    runs = []
    for knobs in ({}, {dsmc.OPT_FUSED: 0}, {dsmc.OPT_MOVE_AHEAD: 0}, {dsmc.OPT_USE_GRAPH: 1},
                  {dsmc.OPT_LDS_CAPACITY: 1024}):
        g = dsmc.System(ini, w)
        g.load_state(rv)
        g.set_sampling(1)
        for k, v in knobs.items():
            g.set_option(k, v)
        g.step(8)
        c = g.counters()
        assert c["errors"] == 0 and c["num_molecules"] == len(rv)
        assert c["overflow_cells"] > 0, knobs         # the dense cell went through the global path
        runs.append((g.state(), g.moments()[0], c["collisions"]))
        g.close()
    st, mom, coll = runs[0]
    assert coll > 1000                                # the dense cell collided
    for s, m, cl in runs[1:]:
        assert np.array_equal(s, st) and np.array_equal(m, mom) and cl == coll
    e1 = np.sum(st[:, 3:] ** 2)
    assert abs(e1 / e0 - 1) < 1e-5
