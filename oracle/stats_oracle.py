"""CPU restatement of the reference's statistics sampler (test infrastructure:
only tests/ import it, as the checker of the device sampler and of the
library's statistics writer).

Follows statisticssampler.cpp:13-48 and statisticalproperty.cpp (line numbers
below) with the unit conversions of unitconverter.cpp:3-60.  Pinned in
tests/test_statistics.py against the compiled reference's own sampler
(oracle/_ref) run on the same states.
"""
import math

import numpy as np

from oracle import binding as B

E0EV = 1.0318e-2          # unitconverter.cpp:8


def _units():
    u = B.units()
    u["E0ev"] = E0EV
    u["P0"] = u["m0"] / (u["t0"] * u["t0"] * u["L0"])        # unitconverter.cpp:14
    u["perm0"] = u["L0"] * u["L0"]                            # 19
    return u


def sample(rv6, mass, temperature, L, bins):
    """One sample of the binned properties and the energy sum from a state
    (absolute positions, reduced units).  Returns dict(count, vz_sum, ez_sum,
    vx_hist, v2_sum)."""
    r, v = rv6[:, :3], rv6[:, 3:]
    # MeasureCount / ...Poiseuille / TemperatureDistribution: 236-240, 266-270, 320-324
    b = (r[:, 1] * (1.0 / L[1]) * bins).astype(np.int64)
    count = np.bincount(b, minlength=bins)[:bins].astype(np.uint64)
    vz_sum = np.bincount(b, weights=v[:, 2], minlength=bins)[:bins]
    ez_sum = np.bincount(b, weights=0.5 * mass * v[:, 2] * v[:, 2], minlength=bins)[:bins]
    # MeasureVelocityDistribution: 398-405 constructor, 413-421 update
    sd = math.sqrt(temperature / mass)
    vmin, vmax = -4 * sd, 4 * sd
    bin_size = (vmax - vmin) / bins
    hb = np.trunc((v[:, 0] - vmin) / bin_size).astype(np.int64)     # C's int conversion
    keep = (hb >= 0) & (hb <= bins - 1)
    hist = np.bincount(hb[keep], minlength=bins)[:bins].astype(np.uint64)
    # MeasureEnergy: 27-29
    v2 = float(np.sum(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]))
    return {"count": count, "vz_sum": vz_sum, "ez_sum": ez_sum, "vx_hist": hist, "v2_sum": v2}


def files(series, sums, p):
    """The statistics files' text.  series: per sample dict(t, v2_sum,
    periodic) with t the system time (reduced) and periodic the flow axis's
    count_periodic; sums: the binned sums over all samples (sample()'s keys);
    p: dict(mass, temperature, fn, n, L, porosity, density, diam, gravity,
    flow_direction, t0)."""
    u = _units()
    mass, fn = p["mass"], p["fn"]
    atoms = p["n"] * int(fn)                                   # system.cpp:509-511
    volume_global = p["L"][0] * p["L"][1] * p["L"][2] * p["porosity"]   # system.cpp:314
    out = {k: [] for k in ("energy", "temperature", "pressure", "number_flow_rate", "volumetric_flow_rate")}
    T = Q = 0.0
    for s in series:
        t_ns = s["t"] * u["t0"] * 1e9                           # system.h:116
        ke = s["v2_sum"] * (0.5 * mass * fn)                    # 29
        out["energy"].append("%f %E\n" % (t_ns, ke * u["E0ev"]))                       # 36
        T = 2.0 / 3.0 * ke / atoms                              # 62
        out["temperature"].append("%f %f\n" % (t_ns, T * u["T0"]))                     # 64
        out["pressure"].append("%f %E\n" % (t_ns, u["P0"] * (atoms / volume_global * T)))   # 156-157
        flux = fn * s["periodic"] / (s["t"] - p["t0"])          # 93
        out["number_flow_rate"].append("%f %E\n" % (t_ns, flux))                      # 97
        Q = flux * (1.0 / p["density"])                         # 124-125
        out["volumetric_flow_rate"].append("%f %E\n" % (t_ns, Q))                     # 127
    text = {k + ".txt": "".join(v) for k, v in out.items()}
    # MeasurePermeability 190-216: the last sample's values
    area = 1.0
    for a in range(3):
        if a != p["flow_direction"]:
            area *= p["L"][a]
    visc = 5.0 / (16 * p["diam"] * p["diam"]) * math.sqrt(mass * T / math.pi)
    k = Q * visc / (area * mass * p["density"] * p["gravity"])
    text["permeability.txt"] = "%E\n" % (k * u["perm0"] * 1e4)
    bins = len(sums["count"])
    c = np.maximum(sums["count"].astype(np.float64), 1.0)
    vel = sums["vz_sum"] / c                                     # 279-289
    tb = sums["ez_sum"] / c * (2.0 / 3.0)                        # 333-341
    pb = sums["count"].astype(np.float64) / (volume_global / bins) * tb     # 377-384
    text["velocity.txt"] = "".join("%E " % (u["v0"] * x) for x in vel)
    text["linear_temperature.txt"] = "".join("%E " % (u["T0"] * x) for x in tb)
    text["linear_pressure.txt"] = "".join("%E " % (u["P0"] * x) for x in pb)
    sd = math.sqrt(p["temperature"] / mass)
    vmin, bin_size = -4 * sd, 8 * sd / bins
    text["velocity_distribution.txt"] = "".join(
        "%f %d\n" % (u["v0"] * (vmin + (i + 0.5) * bin_size), int(sums["vx_hist"][i])) for i in range(bins))
    return text


def numbers(text):
    """The numbers of a statistics file, in order."""
    return np.array([float(x) for x in text.split()])
