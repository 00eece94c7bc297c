"""Warm start: seeding every cell from per-cell fields (density scale, drift,
temperature) instead of the uniform Maxwellian of System::setup_molecules
(system.cpp:398-430)."""
import numpy as np
import pytest

from master_amd import config as C
from tests import dsmc_cases as K


def _expected_counts(types, cpa, N, nrel):
    """Per-cell particle counts the library must produce: shares
    proportional to n_rel x empty voxels, whole particles by cumulative
    rounding, the member total kept."""
    n = types.shape[0]
    b = n // cpa
    empty = (types == 0).reshape(cpa, b, cpa, b, cpa, b).sum(axis=(1, 3, 5)).reshape(-1)
    w = np.where(empty > 0, np.maximum(nrel, 0.0) * empty, 0.0)
    cum = np.cumsum(w)
    upto = np.floor(cum / cum[-1] * N + 0.5).astype(np.int64)
    upto[-1] = N
    return np.diff(np.concatenate([[0], upto])), empty


@pytest.mark.gpu
def test_field_seeding_places_and_draws_as_asked(gpu):
    w = K.world()
    ini = K.settings(w, 200_000, seed=5)
    g = K.gpu_system(ini, w)
    nc = g.ncells
    cpa = int(ini["cells_x"])
    T, m, _ = C.gas_scales(ini)
    rng = np.random.default_rng(0)
    f = np.zeros((nc, 5))
    f[:, 0] = rng.uniform(0.5, 1.5, nc)
    f[:, 1] = 0.05 * np.sin(np.arange(nc) / 7.0)
    f[:, 2] = -0.03
    f[:, 3] = 0.02 * np.cos(np.arange(nc) / 5.0)
    f[:, 4] = T * rng.uniform(0.7, 1.3, nc)
    n = g.initialize_fields(f)
    c = g.counters()
    assert c["errors"] == 0
    types = np.asarray(w["types"]).reshape(w["n"])
    ref_n = K.gpu_system(ini, w).initialize()          # what the uniform seeding gives
    assert n == ref_n
    want, empty = _expected_counts(types, cpa, n, f[:, 0])
    st = g.state()
    cell = K.cell_of_positions(st, ini)
    assert np.all(np.diff(cell) >= 0)                   # sorted
    got = np.bincount(cell, minlength=nc)
    assert np.array_equal(got, want)
    # only empty voxels
    h = float(ini["Lx"]) / w["n"][0]
    vox = np.floor(st[:, :3] / h).astype(np.int64)
    assert np.all(types[vox[:, 0], vox[:, 1], vox[:, 2]] == 0)
    # per-cell drift and temperature within 5 standard errors
    act = got >= 50
    v = st[:, 3:]
    for k in range(3):
        mean = np.bincount(cell, weights=v[:, k], minlength=nc)[act] / got[act]
        se = np.sqrt(f[act, 4] / m / got[act])
        assert np.all(np.abs(mean - f[act, 1 + k]) < 5 * se + 1e-6), k
    var = sum(np.bincount(cell, weights=v[:, k] ** 2, minlength=nc)[act] / got[act]
              - (np.bincount(cell, weights=v[:, k], minlength=nc)[act] / got[act]) ** 2 for k in range(3)) / 3
    Tm = m * var
    rel_se = np.sqrt(2.0 / (3 * got[act]))
    assert np.mean(np.abs(Tm / f[act, 4] - 1) / rel_se) < 1.2
    assert np.all(np.abs(Tm / f[act, 4] - 1) < 6 * rel_se)
    # and it steps
    g.step(10)
    assert g.counters()["errors"] == 0
