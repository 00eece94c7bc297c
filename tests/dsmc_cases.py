"""Shared workload builders for the parity tests (small enough for the oracle)."""
import numpy as np

from master_amd import config as C
from master_amd import dsmc
from oracle import binding as B

_WORLDS = {}


def world(n=32, r=4.0, phi=0.6, seed=1):
    key = (n, r, phi, seed)
    if key not in _WORLDS:
        _WORLDS[key] = B.packed_spheres(n, r, phi, seed)
    return _WORLDS[key]


def settings(w, n_particles, seed=1, wall="thermal", gravity=None, kn=0.1, cells=None, dt_tau=0.2):
    n = w["n"][0]
    ini = C.dsmc_settings(n, w["porosity"], n_particles, kn, cells or n // 4, seed=seed, wall=wall,
                          dt_tau=dt_tau)
    if gravity is not None:
        ini["gravity"] = gravity
    return ini


def gpu_system(ini, w):
    return dsmc.System(ini, w)


def oracle_system(ini, w, state=None):
    o = B.OracleSystem(B.params_from_settings(ini, w, rng_mode=B.PHILOX), w)
    if state is None:
        o.seed()
    else:
        o.load_state(state)
    return o


def cell_of_positions(rv6, ini):
    c = int(ini["cells_x"])
    L = float(ini["Lx"])
    ijk = np.floor(rv6[:, :3] / L * c).astype(np.int64)
    return (ijk[:, 0] * c + ijk[:, 1]) * c + ijk[:, 2]


def vel_err(a, b, ini):
    """|a - b| / max(|b|, sqrt(T/m)): relative error with the thermal speed
    as the floor (a slow particle's velocity carries the rounding of the
    larger speeds it was computed from)."""
    T, m, _ = C.gas_scales(ini)
    floor = np.sqrt(T / m)
    d = np.linalg.norm(a - b, axis=1)
    return d / np.maximum(np.linalg.norm(b, axis=1), floor)


def rel_vec_err(a, b):
    """|a - b| / max(|b|, tiny) per row."""
    d = np.linalg.norm(a - b, axis=1)
    s = np.maximum(np.linalg.norm(b, axis=1), 1e-30)
    return d / s
