"""The movie writer (DSMC_IO::save_state_to_movie_file, dsmc_io.cpp:50-80,
called before every step by main.cpp:57-59) and the run keys of dsmc.ini
(settings.cpp:6-14).

The CPU restatement (oracle/movie_oracle.py) is pinned byte for byte against
the compiled reference's own writer on the reference's states; the device
frames are then checked against the library's own states: every row is a
particle of the state at that step (the same float32 of the FP64 position
dsmc_get_state returns), the rows follow the same particles from frame to
frame (free flights of v dt), the first frame holds the loaded file's first
molecules, and the trajectory is the one without a movie."""
import numpy as np
import pytest

from master_amd import config as C
from oracle import binding as B
from oracle import movie_oracle as M
from tests import dsmc_cases as K


def _ref_count(ini, w):
    return B.RefSystem(ini, w).n


@pytest.mark.skipif(not B.ref_available(), reason="reference harness not built")
@pytest.mark.parametrize("every,over", [(2, False), (1, True)])
def test_restatement_reproduces_the_reference_writer(every, over):
    """Frames before steps 0..5 of the reference's run: count below the
    number of molecules, and above it (rows of the zeroed buffer)."""
    w = K.world()
    ini = K.settings(w, 3000, gravity=0.5)
    n = _ref_count(ini, w)
    count = n + 37 if over else n // 3
    ini["create_movie"], ini["movie_every_n_frame"], ini["movie_molecules"] = "true", every, count
    r = B.RefSystem(ini, w)
    mine = M.MovieWriter(count, int(ini["max_molecules_per_node"]))
    data = b""
    for k in range(6):
        data = r.movie_frame()
        mine.before_step(k, every, r.state())
        r.step()
    assert bytes(mine.data) == data
    frames = M.read_frames(data)
    assert len(frames) == len(range(0, 6, every)) and all(f.shape == (count, 6) for f in frames)
    if over:
        assert not frames[-1][n:].any() and frames[-1][:n].any()


def test_run_settings_from_ini(tmp_path):
    from master_amd import dsmc
    ini = dict(C.BASE_INI)
    ini.update({"timesteps": 250, "create_movie": "true", "movie_every_n_frame": 7, "movie_molecules": 4096,
                "statistics_interval": 25, "sampling_bins": 64, "load_previous_state": "True"})
    B.write_ini(tmp_path / "dsmc.ini", ini)
    r = dsmc.run_settings_from_ini_file(tmp_path / "dsmc.ini")
    # CIniFile::getbool: only the text "true" is true (cinifile.h:73-82)
    assert r == {"timesteps": 250, "load_previous_state": 0, "create_movie": 1, "movie_every_n_frame": 7,
                 "movie_molecules": 4096, "statistics_interval": 25, "sampling_bins": 64}
    del ini["movie_molecules"]
    B.write_ini(tmp_path / "bad.ini", ini)
    with pytest.raises(dsmc.DsmcError):
        dsmc.run_settings_from_ini_file(tmp_path / "bad.ini")


def _rows(a):
    return {r.tobytes() for r in np.ascontiguousarray(a, np.float32)}


@pytest.mark.gpu
def test_frames_follow_the_loaded_order(gpu, tmp_path):
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000, gravity=0.5)
    g0 = dsmc.System(ini, w)
    g0.initialize()
    g0.step(3)
    rv = g0.state()
    g0.close()
    rv = rv[np.random.default_rng(5).permutation(len(rv))]     # a file order unlike the cell order
    m = 1500
    g = dsmc.System(ini, w)
    g.set_movie(tmp_path / "movie.bin", 3, m)
    g.load_state(rv)
    frames = []
    for _ in range(3):
        f = g.movie_frame()
        rows = _rows(g.state())
        assert all(r.tobytes() in rows for r in f)      # each row a particle of this step's state
        assert len(_rows(f)) == m                       # distinct particles
        frames.append(f)
        g.step(3)                                       # the frame of this step, then three steps
    g.set_movie(None)
    got = M.read_frames(open(tmp_path / "movie.bin", "rb").read())
    assert len(got) == 3 and all(np.array_equal(a, b) for a, b in zip(got, frames))
    # the first frame: the file's first m molecules (positions through the
    # voxel + float offset form of the state)
    L = float(ini["Lx"])
    assert np.array_equal(frames[0][:, 3:], rv[:m, 3:].astype(np.float32))
    assert np.abs(frames[0][:, :3] - rv[:m, :3]).max() < 1e-6 * L
    g.close()


@pytest.mark.gpu
def test_rows_track_particles_between_frames(gpu, tmp_path):
    """No gravity: between frames a particle that meets no wall flies v dt
    (periodic wrap aside).  Rows of other particles would not."""
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000, gravity=0.0)
    g = dsmc.System(ini, w)
    g.initialize()
    m = 4000
    g.set_movie(tmp_path / "movie.bin", 1, m)
    st = g.state().astype(np.float32)
    f0 = g.movie_frame()
    assert np.array_equal(f0, st[:m])                   # set after initialize: the current order
    g.step(4)
    g.set_movie(None)
    frames = M.read_frames(open(tmp_path / "movie.bin", "rb").read())
    assert len(frames) == 4 and np.array_equal(frames[0], f0)
    L, dt = float(ini["Lx"]), float(ini["dt"])
    for a, b in zip(frames[:-1], frames[1:]):
        d = b[:, :3].astype(np.float64) - a[:, :3]
        d -= L * np.round(d / L)
        err = np.abs(d - a[:, 3:].astype(np.float64) * dt).max(axis=1)
        assert np.mean(err < 1e-5 * L) > 0.8


@pytest.mark.gpu
def test_movie_leaves_the_trajectory_unchanged(gpu, tmp_path):
    """Graphs and eager steps, with the sort carrying ids, give the trajectory
    of a run without a movie and the same frames whatever the launch options
    (a movie sets moving ahead aside); an ensemble's rows come from the
    member holding each particle."""
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000, gravity=0.5)
    runs = {}
    for name, movie, graph, ahead in [("plain", False, True, 1), ("graph", True, True, 1),
                                      ("eager", True, False, 1), ("noahead", True, False, 0)]:
        g = dsmc.System(ini, w)
        g.initialize()
        g.set_option(dsmc.OPT_USE_GRAPH, int(graph))
        g.set_option(dsmc.OPT_MOVE_AHEAD, ahead)
        if graph:
            g.set_option(dsmc.OPT_GRAPH_STEPS, 4)
        if movie:
            g.set_movie(tmp_path / f"{name}.bin", 4, 700)
        g.step(12)
        runs[name] = g.state()
        g.close()
    for name in ("graph", "eager", "noahead"):
        assert np.array_equal(runs[name], runs["plain"]), name
    ref = M.read_frames(open(tmp_path / "noahead.bin", "rb").read())
    assert len(ref) == 3
    for name in ("graph", "eager"):
        got = M.read_frames(open(tmp_path / f"{name}.bin", "rb").read())
        assert len(got) == 3 and all(np.array_equal(x, y) for x, y in zip(got, ref)), name

    w2 = K.world(n=32, r=3.0, phi=0.7, seed=4)
    ini2 = K.settings(w2, 15000, seed=9)
    e = dsmc.System([ini, ini2], [w, w2])
    e.initialize()
    e.set_movie(tmp_path / "ens.bin", 1, 60000)         # more rows than particles: the tail stays zero
    e.step(3)
    f = e.movie_frame()
    n = e.n
    rows = _rows(e.state())
    assert all(r.tobytes() in rows for r in f[:n]) and not f[n:].any()
    e.close()
