"""Settings for the DSMC step in the reference's own vocabulary.

`settings.cpp:3-44` reads a flat ``key = value`` file (dsmc.ini) with these
keys; `System::initialize` (system.cpp:253-266) converts them to reduced units
through `UnitConverter` (unitconverter.cpp:3-21: m0 = 1 amu, L0 = 1 um,
E0 = 1.65088e-21 J).  This module builds such key/value dictionaries for the
workloads this library is measured on, from physical inputs (Knudsen number
against the voxel edge, particle count, porosity), using the closed forms of
the reference's job driver (dsmc_unit_converter.py:112-116: n = 1/(sqrt(2) pi
d^2 lambda)).
"""
from __future__ import annotations

import math

# unitconverter.cpp:5-19
M0 = 1.66053886e-27
L0 = 1e-6
E0 = 1.65088e-21
KB = 1.3806503e-23
T0_TIME = L0 * math.sqrt(M0 / E0)
T0_TEMP = E0 / KB
V0 = L0 / T0_TIME
A0 = V0 / T0_TIME
NUMBER_DENSITY0 = 1.0 / (L0 * L0 * L0)

ARGON_MASS_KG = 6.63e-26
ARGON_DIAM_M = 3.66e-10

# Every key Settings::Settings reads (settings.cpp:6-42) plus `world`
# (system.cpp:290).  Keys the step does not use get the job driver's values.
BASE_INI = {
    "load_previous_state": "false",
    "create_movie": "false",
    "max_molecules_per_node": 0,
    "atoms_per_molecule": 1,
    "timesteps": 1000,
    "temperature": 300.0,
    "wall_temperature": 300.0,
    "statistics_interval": 100,
    "velocity_profile_type": "other",
    "sampling_bins": 100,
    "movie_every_n_frame": 1,
    "movie_molecules": 0,
    "dt": 0.001,
    "gravity": 0.0,
    "flow_direction": 0,
    "pressure_A": 0.0,
    "pressure_B": 0.0,
    "reservoir_fraction": 0.0,
    "world": "world",
    "surface_interaction_model": "thermal",
    "alpha_n": 1.0,
    "alpha_t": 1.0,
    "seed": 1,
    "density": 1.0,
    "diam": ARGON_DIAM_M / L0,
    "mass": ARGON_MASS_KG / M0,
    "viscosity": 2.27e-5,
    "Lx": 1.0, "Ly": 1.0, "Lz": 1.0,
    "cells_x": 8, "cells_y": 8, "cells_z": 8,
    "nx": 1, "ny": 1, "nz": 1,
}


def gas_scales(ini):
    """Reduced mean free path helpers for a settings dict."""
    T = float(ini["temperature"]) / T0_TEMP
    m = float(ini["mass"])
    cbar = math.sqrt(8.0 * T / (math.pi * m))
    return T, m, cbar


def dsmc_settings(n_vox, porosity, n_particles, kn, cells_per_axis, seed=1, dt_tau=0.2,
                  voxel_edge_um=1.0 / 32.0, gravity=None, mach=0.05, wall="thermal", timesteps=1000):
    """Settings for an n_vox^3 periodic porous box.

    Kn is against the voxel edge: lambda = kn * h.  Number density from
    lambda = 1/(sqrt(2) pi d^2 n); dt = dt_tau * lambda / cbar; Fn (the
    reference's int atoms_per_molecule) from N = n V phi / Fn.  The body force
    defaults to a slip-Poiseuille estimate for a drift of `mach` times the
    speed of sound in pores of ~5 voxels.
    """
    ini = dict(BASE_INI)
    L = n_vox * voxel_edge_um
    h = voxel_edge_um
    d = ARGON_DIAM_M / L0
    lam = kn * h
    n_red = 1.0 / (math.sqrt(2.0) * math.pi * d * d * lam)
    ini.update({
        "Lx": L, "Ly": L, "Lz": L,
        "cells_x": cells_per_axis, "cells_y": cells_per_axis, "cells_z": cells_per_axis,
        "density": n_red * NUMBER_DENSITY0,
        "seed": seed,
        "surface_interaction_model": wall,
        "timesteps": timesteps,
    })
    T, m, cbar = gas_scales(ini)
    ini["dt"] = dt_tau * lam / cbar
    fn = max(1, int(round(n_red * L ** 3 * porosity / n_particles)))
    ini["atoms_per_molecule"] = fn
    ini["max_molecules_per_node"] = int(n_particles * 1.5) + 1000
    if gravity is None:
        # speed of sound for a monatomic gas: sqrt(5/3 kT/m); reduced units
        c_s = math.sqrt(5.0 / 3.0 * T / m)
        u = mach * c_s
        # kinematic viscosity of a hard-sphere gas, Chapman-Enskog first order:
        # mu = 5/16 sqrt(pi m k T)/(pi d^2)
        mu = 5.0 / 16.0 * math.sqrt(math.pi * m * T) / (math.pi * d * d)
        nu = mu / (n_red * m)
        H = 5.0 * h
        gravity = 12.0 * nu * u / (H * H)
    ini["gravity"] = gravity
    ini["flow_direction"] = 0
    return ini


# ------------------------------------------------------------------ workloads
# Plain DSMC, CPU-runnable size: 32^3, spheres r=4 voxels, phi target 0.6,
# seed 1, 8^3 cells of 4^3 voxels, 200k particles, Kn 0.1, dt 0.2 tau.
CPU_SIZE = dict(n_vox=32, radius_vox=4.0, porosity=0.6, geom_seed=1, n_particles=200_000,
                kn=0.1, cells=8)
# Precision form: 64^3, r=6, phi 0.5, seed 2, 16^3 cells, 2M particles.
PRECISION = dict(n_vox=64, radius_vox=6.0, porosity=0.5, geom_seed=2, n_particles=2_000_000,
                 kn=0.1, cells=16)
# Roofline run: 256^3, radii log-uniform 3..10 voxels, phi 0.45, seed 3,
# 64^3 cells, 100M particles, Kn 0.05.
ROOFLINE = dict(n_vox=256, rmin=3.0, rmax=10.0, porosity=0.45, geom_seed=3,
                n_particles=100_000_000, kn=0.05, cells=64)


# Batched ensemble: 64 independent porous boxes advanced by one handle.
# Edge from {32, 48, 64, 96, 128} voxels, porosity U(0.3, 0.8), sphere radius
# U(2, 8) voxels, Kn log-uniform in [0.01, 1], body force for a drift of
# Mach log-uniform in [0.01, 0.1], ~8 particles per pore voxel (clamped to
# 1e5..8e6 per member), seeds 100..163.  Deterministic in (n_members, seed0).
ENSEMBLE_EDGES = (32, 48, 64, 96, 128)
ENSEMBLE_MIN_N, ENSEMBLE_MAX_N = 100_000, 8_000_000     # particles per member


def ensemble_members(n_members=64, seed0=100, per_voxel=8.0, edges=ENSEMBLE_EDGES):
    import numpy as np
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(n_members):
        out.append(dict(
            edge=int(rng.choice(edges)),
            porosity=float(rng.uniform(0.3, 0.8)),
            radius=float(rng.uniform(2.0, 8.0)),
            kn=float(10.0 ** rng.uniform(-2.0, 0.0)),
            mach=float(10.0 ** rng.uniform(-2.0, -1.0)),
            seed=seed0 + i,
            per_voxel=per_voxel,
        ))
    return out


def member_settings(spec, porosity):
    """Settings of one ensemble member for its generated world's porosity."""
    n = spec["edge"]
    n_particles = int(min(max(spec["per_voxel"] * n ** 3 * porosity, ENSEMBLE_MIN_N), ENSEMBLE_MAX_N))
    return dsmc_settings(n, porosity, n_particles, spec["kn"], n // 4, seed=spec["seed"], mach=spec["mach"])
