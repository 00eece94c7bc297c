"""Cercignani-Lampis walls (collidercercignanilampis.cpp:13-45 with
Random::next_cercignani_lampis_normal_component, random.cpp:27-91), pinned on
the CPU:

* the accepted distribution of the reference's loop in closed form: proposals
  N(mu, alpha_n / 2), mu = sqrt(1 - alpha_n) w_in, kept with probability
  min(1, b(v)), b = v exp(-arg) I0(arg) / b_max, so p(v) is the normal
  density times min(1, b) on v > 0 (velocities in units of
  sqrt(2 T_wall / m)); I0 from scipy, not from the polynomials under test;
* the compiled reference's own loop (oracle/_ref) against that closed form,
  and the oracle's ran1 restatement of it draw for draw;
* the oracle's truncated-proposal sampler (the device's algorithm) against the
  closed form, including fast incoming particles where the reference's loop
  would need ~1e8 proposals;
* tangential accommodation from the oracle's wall log: alpha_t = 0 keeps both
  tangential components, alpha_t = 1 draws them from N(0, T_wall / m), and in
  between the outgoing component regresses on the incoming one with slope
  sqrt(1 - alpha_t) and residual variance alpha_t T_wall / m.
"""
import numpy as np
import pytest
from scipy import special, stats

from oracle import binding as B
from tests import dsmc_cases as K

BMAX = 4.25     # v_max = 1500 m/s over sqrt(2 k 300 K / 40 amu)


def accepted_cdf(alpha_n, w_in, bmax=BMAX):
    mu, sd = np.sqrt(1 - alpha_n) * w_in, np.sqrt(alpha_n / 2)
    v = np.linspace(1e-12, max(mu, 0.0) + 14 * sd + 6.0, 400001)
    arg = 2 * np.sqrt(1 - alpha_n) / alpha_n * w_in * v
    e_i0 = special.i0e(np.abs(arg)) * np.exp(np.abs(arg) - arg)        # exp(-arg) I0(arg)
    pdf = stats.norm.pdf(v, mu, sd) * np.minimum(1.0, v * e_i0 / bmax)
    c = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(v))])
    c /= c[-1]
    return lambda x: np.interp(x, v, c)


def test_bessel_polynomials_match_scipy():
    """Random::bessel_i0's polynomials (A&S 9.8.1/9.8.2, |rel. error| <
    1.9e-7) in log(exp(-x) I0(x)) form, against scipy's i0e on both sides of
    the 3.75 switch and far into the range fast particles reach."""
    xs = np.concatenate([np.linspace(-60, 60, 2401), [-3.75, 3.75, -3.7499, 3.7501, 0.0]])
    got = np.array([B.cl_log_exp_i0(x) for x in xs])
    exact = np.log(special.i0e(np.abs(xs))) + np.abs(xs) - xs
    assert np.max(np.abs(got - exact)) < 5e-7


@pytest.mark.skipif(not B.ref_available(), reason="reference harness not built")
def test_oracle_ran1_loop_is_the_reference_loop():
    for alpha_n, w_in, seed in [(0.5, -0.3, -5), (0.9, 0.4, -11), (1.0, -0.2, -3)]:
        a, ta = B.cl_normal_samples(alpha_n, w_in, BMAX, 3000, rng_mode=B.RAN1, seed=seed)
        r, tr = B.cl_normal_samples(alpha_n, w_in, BMAX, 3000, use_ref=True, seed=seed)
        assert np.array_equal(a, r) and ta == tr


@pytest.mark.skipif(not B.ref_available(), reason="reference harness not built")
@pytest.mark.parametrize("alpha_n,w_in", [(0.5, -0.3), (1.0, -0.5), (0.8, 0.0), (0.5, 1.0), (0.3, -0.6)])
def test_reference_loop_has_the_closed_form_distribution(alpha_n, w_in):
    s, _ = B.cl_normal_samples(alpha_n, w_in, BMAX, 20000, use_ref=True, seed=-7)
    assert stats.kstest(s, accepted_cdf(alpha_n, w_in)).pvalue > 1e-3


@pytest.mark.parametrize("alpha_n,w_in", [(0.5, -0.3), (1.0, -0.5), (0.2, -2.5), (0.8, 0.0), (0.5, -4.0),
                                          (0.05, -1.0), (0.5, 1.0), (0.3, -0.6)])
def test_truncated_sampler_has_the_closed_form_distribution(alpha_n, w_in):
    s, trials = B.cl_normal_samples(alpha_n, w_in, BMAX, 20000, key=3, step=2)
    assert np.all(s > 0)
    assert stats.kstest(s, accepted_cdf(alpha_n, w_in)).pvalue > 1e-3
    assert trials < 40 * len(s)          # the plain loop needs ~1e8 per sample at w_in = -4


def test_full_normal_accommodation_is_rayleigh():
    """alpha_n = 1: b = v / b_max, the outgoing normal speed is Rayleigh,
    F(v) = 1 - exp(-v^2), as the thermal wall's (colliderthermal.cpp)."""
    s, _ = B.cl_normal_samples(1.0, -1.7, BMAX, 20000, key=5)
    assert stats.kstest(s, lambda x: 1 - np.exp(-x * x)).pvalue > 1e-3


def _wall_hits(alpha_n, alpha_t, steps=3):
    w = K.world()
    ini = K.settings(w, 20000, wall="cercignani_lampis", dt_tau=2.0)
    ini["alpha_n"], ini["alpha_t"] = alpha_n, alpha_t
    o = K.oracle_system(ini, w)
    o.enable_wall_log(2_000_000)
    vin, vout, vox = [], [], []
    for _ in range(steps):
        o.step(1)
        g, _, a, b = o.wall_log()
        vin.append(a), vout.append(b), vox.append(g)
    assert o.counters()["errors"] == 0
    g = np.concatenate(vox)
    frame = [np.asarray(w[k], np.float64).reshape(-1, 3)[g] for k in ("normals", "t1", "t2")]
    sw = np.sqrt(float(ini["wall_temperature"]) / B.units()["T0"] / float(ini["mass"]))
    return np.concatenate(vin), np.concatenate(vout), frame, sw


@pytest.mark.parametrize("alpha_t", [0.0, 0.4, 1.0])
def test_tangential_accommodation(alpha_t):
    vin, vout, (n, t1, t2), sw = _wall_hits(0.6, alpha_t)
    assert len(vin) > 2000
    assert np.all(np.sum(vout * n, 1) > 0)                 # every particle leaves the wall
    for t in (t1, t2):
        a, b = np.sum(vin * t, 1), np.sum(vout * t, 1)
        if alpha_t == 0.0:
            # kept up to the float frames' orthogonality (|n . t| ~ 2e-6 here)
            assert np.allclose(b, a, rtol=0, atol=1e-5 * sw)
            continue
        slope, icpt = np.polyfit(a, b, 1)
        resid = b - slope * a - icpt
        se = np.sqrt(alpha_t) * sw / (np.std(a) * np.sqrt(len(a)))
        assert abs(slope - np.sqrt(1 - alpha_t)) < 5 * se
        assert abs(np.std(resid) / (np.sqrt(alpha_t) * sw) - 1) < 5 / np.sqrt(2 * len(a))
        if alpha_t == 1.0:
            assert stats.kstest(b / sw, "norm").pvalue > 1e-3
