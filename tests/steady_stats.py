"""Per-cell steady-state fields (density, drift velocity, temperature) from
windowed moment sums, and the ensemble z-test used to compare two sets of
independent runs (different seeds) of the same configuration.

Fields per cell over the steady part of a run (windows k0..end):
  n  = particles per step, u = sum(v) / sum(1), T = m (<v^2> - <v>^2) / 3
(matlab/3d/calculatePressure.m:11-32 for the per-cell moment form).

The variance of a run's per-cell mean is estimated from the spread across
runs (seeds), pooled over cells of similar occupancy: cell densities relax
by diffusion over a cell edge, which takes longer than a run, so windows of
one run are not independent samples of the density.
"""
import numpy as np

FIELDS = ("n", "ux", "uy", "uz", "T")


def run_fields(sums, mass, k0):
    """sums: (nwin, ncell, 7) windowed sums of (1, v, v^2); returns
    (ncell, 5) fields over windows k0..end and the number of steps."""
    s = np.asarray(sums)[k0:].sum(0)
    n = s[:, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        u = s[:, 1:4] / n[:, None]
        var = s[:, 4:7] / n[:, None] - u ** 2
        T = mass * var.sum(-1) / 3.0
    return np.column_stack([n, u, T])


def ensemble(runs_sums, mass, k0, steps_per_window):
    """Stacks the fields of several runs: (nruns, ncell, 5), n per step."""
    F = np.stack([run_fields(s, mass, k0) for s in runs_sums])
    nsteps = (np.asarray(runs_sums[0]).shape[0] - k0) * steps_per_window
    F[..., 0] /= nsteps
    return F


def pooled_variance(FA, FB, active, neighbours=32):
    """Per-cell variance of one run's field, from the spread across runs of
    both ensembles (each about its own mean), smoothed over the `neighbours`
    cells of closest mean occupancy."""
    na, nb = len(FA), len(FB)
    va = FA.var(0, ddof=1) * (na - 1)
    vb = FB.var(0, ddof=1) * (nb - 1)
    v = (va + vb) / (na + nb - 2)                     # (ncell, 5)
    occ = np.concatenate([FA[..., 0], FB[..., 0]]).mean(0)
    idx = np.nonzero(active)[0]
    order = idx[np.argsort(occ[idx])]
    out = np.full_like(v, np.nan)
    h = neighbours // 2
    for r, c in enumerate(order):
        lo, hi = max(0, r - h), min(len(order), r + h + 1)
        out[c] = v[order[lo:hi]].mean(0)
    return out


def z_scores(FA, FB, min_particles=5.0):
    """z per active cell and field for mean(FA) - mean(FB)."""
    active = (FA[..., 0].min(0) > min_particles) & (FB[..., 0].min(0) > min_particles)
    var = pooled_variance(FA, FB, active)
    se = np.sqrt(var * (1.0 / len(FA) + 1.0 / len(FB)))
    z = (FA.mean(0) - FB.mean(0)) / se
    return z[active], active


def summarize(z):
    return {f: {"mean_z2": float(np.mean(z[:, i] ** 2)), "max_abs": float(np.max(np.abs(z[:, i]))),
                "frac_gt_3p5": float(np.mean(np.abs(z[:, i]) > 3.5))} for i, f in enumerate(FIELDS)}


def consistent(z, mean_z2_max=1.6, max_abs=5.0, frac_max=0.01):
    s = summarize(z)
    bad = {f: v for f, v in s.items()
           if v["mean_z2"] > mean_z2_max or v["max_abs"] > max_abs or v["frac_gt_3p5"] > frac_max}
    return not bad, s
