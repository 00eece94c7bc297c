"""dsmc_run, the reference's main program (main.cpp:19-103) over the C ABI:
a reference run directory in, the reference's output files out (state file,
movie, statistics, Tocontinue); a resumed run (load_previous_state with
Tocontinue) continues the step counter, so it equals one long run of the
library from the saved state."""
import os
import subprocess

import numpy as np
import pytest

from master_amd import dsmc as DS
from oracle import binding as B
from oracle import movie_oracle as M
from tests import dsmc_cases as K

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "master_amd", "dsmc_run")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="dsmc_run is not built")


def _run(d):
    print(f"[dsmc_run] {d}", flush=True)
    return subprocess.run([DRIVER, str(d)], capture_output=True, text=True, timeout=300)


def _state_file(path):
    """state_files/stateNNNN.bin: int N, then N x (r, v) doubles (dsmc_io.cpp:82-119)."""
    data = open(path, "rb").read()
    n = int(np.frombuffer(data, np.int32, 1)[0])
    return np.frombuffer(data, np.float64, 6 * n, 4).reshape(n, 6)


@needs_driver
def test_driver_reports_a_missing_run_directory_and_ini(tmp_path):
    r = _run(tmp_path / "missing")
    assert r.returncode == 2
    r = _run(tmp_path)
    assert r.returncode == 1 and "dsmc.ini" in r.stderr


@pytest.mark.gpu
@needs_driver
def test_driver_runs_and_resumes(gpu, tmp_path):
    dsmc = gpu
    w = K.world()
    ini = K.settings(w, 20000, gravity=0.5)
    ini.update({"timesteps": 20, "create_movie": "true", "movie_every_n_frame": 5, "movie_molecules": 300,
                "statistics_interval": 5, "sampling_bins": 10, "load_previous_state": "false", "world": "world"})
    B.write_ini(tmp_path / "dsmc.ini", ini)
    os.makedirs(tmp_path / "world")
    DS.WorldData.from_dict(w).write(str(tmp_path / "world" / "0000.bin"))
    r = _run(tmp_path)
    assert r.returncode == 0, r.stderr
    t, steps, coll, wall = open(tmp_path / "Tocontinue").read().split()
    assert int(steps) == 20
    st = _state_file(tmp_path / "state_files" / "state0000.bin")
    g = dsmc.System(ini, w)
    g.initialize()
    g.step(20)
    assert np.array_equal(st, g.state())
    c = g.counters()
    assert int(coll) == c["collisions"] and int(wall) == c["wall_collisions"]
    frames = M.read_frames(open(tmp_path / "movie_files" / "movie0000.bin", "rb").read())
    assert len(frames) == 4 and all(f.shape == (300, 6) for f in frames)     # steps 0, 5, 10, 15
    assert len(open(tmp_path / "statistics" / "energy.txt").read().splitlines()) == 4   # after 5, 10, 15, 20
    g.close()

    ini.update({"load_previous_state": "true", "timesteps": 10})
    B.write_ini(tmp_path / "dsmc.ini", ini)
    r = _run(tmp_path)
    assert r.returncode == 0, r.stderr
    t2, steps2, _, _ = open(tmp_path / "Tocontinue").read().split()
    assert int(steps2) == 30 and float(t2) > float(t)
    st2 = _state_file(tmp_path / "state_files" / "state0000.bin")
    g2 = dsmc.System(ini, w)
    g2.load_state(st)
    g2.set_counters(20, float(t), int(coll), int(wall))
    g2.step(10)
    assert np.array_equal(st2, g2.state())
    frames2 = M.read_frames(open(tmp_path / "movie_files" / "movie0000.bin", "rb").read())
    assert len(frames2) == 2                            # steps 20 and 25
    # the loaded state's first molecules, through the voxel + offset form
    assert np.array_equal(frames2[0][:, 3:], st[:300, 3:].astype(np.float32))
    assert np.abs(frames2[0][:, :3] - st[:300, :3]).max() < 1e-6 * float(ini["Lx"])
    assert len(open(tmp_path / "statistics" / "energy.txt").read().splitlines()) == 2
    g2.close()
