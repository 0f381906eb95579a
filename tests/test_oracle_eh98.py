"""Pins for the oracle's Eisenstein & Hu (1998) transfer function and the
sigma8-normalised linear power spectrum (oracle/cosmology.py), from what
physics and mathematics fix independently of the fitting formulae' text:

* the coefficients of eqs. 3 and 5 from CODATA constants (k_eq from z_eq,
  R from the photon density of a 2.7 K black body);
* the closed-form sound horizon (eq. 6) against a direct quadrature of the
  sound-horizon integral it solves, and against the paper's own fitting
  form (eq. 26);
* the zero-baryon limit, where eq. 16 reduces to T0(k, 1, 1);
* the nodes of the baryon transfer function at k s~(k) = m pi (eq. 21);
* agreement with the paper's no-wiggle form (eqs. 26-31), which fits the
  same physics with other formulae;
* sigma8 from an independent quadrature of the normalised P(k), the
  primordial slope at k -> 0 and the growth scaling with a.

A mutation test plants plausible typos into a copy of the oracle's EH98 code
(on the CPU) and checks that at least one pin fails for each.

jax_cosmo (the reference's source of these numbers, jaxpm/pm.py:2) cannot be
imported here; it implements the same published equations.
"""
import inspect
import textwrap

import numpy as np
import pytest
from scipy import constants as C
from scipy.integrate import quad
from scipy.optimize import brentq

from oracle import cosmology as oc

FIDUCIAL = oc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
COSMOS = [FIDUCIAL, oc.Cosmology(), oc.Cosmology(Omega_c=0.2, Omega_b=0.04, h=0.65, n_s=1.0, sigma8=0.9),
          oc.Cosmology(Omega_c=0.35, Omega_b=0.06, h=0.75)]


def _photon_omega_h2(T):
    """Omega_gamma h^2 of a black body at T kelvin (CODATA constants)."""
    H0 = 100e3 / (C.parsec * 1e6)  # 100 km/s/Mpc in 1/s
    rho_crit = 3 * H0 ** 2 / (8 * np.pi * C.G)
    return 4 * C.sigma * T ** 4 / C.c ** 3 / rho_crit


def _no_wiggle(c, k):
    """EH98's no-wiggle transfer function (eqs. 26-31), k in h/Mpc."""
    wm = c.Omega_m * c.h ** 2
    wb = c.Omega_b * c.h ** 2
    th = oc.T_CMB / 2.7
    fb = c.Omega_b / c.Omega_m
    s = 44.5 * np.log(9.83 / wm) / np.sqrt(1 + 10 * wb ** 0.75)
    ag = 1 - 0.328 * np.log(431 * wm) * fb + 0.38 * np.log(22.3 * wm) * fb ** 2
    gam = c.Omega_m * c.h * (ag + (1 - ag) / (1 + (0.43 * k * c.h * s) ** 4))
    qq = k * th ** 2 / gam
    L0 = np.log(2 * np.e + 1.8 * qq)
    return L0 / (L0 + (14.2 + 731 / (1 + 62.5 * qq)) * qq * qq)


def _pins(parts_fn, lmp_fn=None):
    """Every pin below, as one callable over a parts function (so the
    mutation test can run them against a planted mistake)."""
    for c in COSMOS:
        wm = c.Omega_m * c.h ** 2
        wb = c.Omega_b * c.h ** 2
        th = oc.T_CMB / 2.7
        p = parts_fn(c, np.array([0.1]))
        # eq. 3: k_eq = sqrt(2 Omega_0 z_eq) H0 / c
        keq = np.sqrt(2 * c.Omega_m * p["zeq"]) * c.h / (C.c / 1e5)
        assert abs(p["keq"] / keq - 1) < 5e-4, ("keq", p["keq"], keq)
        # eq. 2: z_eq = rho_m / rho_r with three massless neutrinos
        wr = _photon_omega_h2(2.7) * th ** 4 * (1 + 3 * 7 / 8 * (4 / 11) ** (4 / 3))
        assert abs(p["zeq"] / (wm / wr) - 1) < 2e-3, ("zeq", p["zeq"], wm / wr)
        # eq. 5: R = 3 rho_b / (4 rho_gamma) at z (EH98 writes z for 1 + z)
        for z, R in ((p["zd"], p["Rd"]), (p["zeq"], p["Req"])):
            Rphys = 0.75 * wb / (_photon_omega_h2(2.7) * th ** 4) / z
            assert abs(R / Rphys - 1) < 1e-3, ("R", R, Rphys)
        # eq. 6 is the integral of c_s / (a^2 H) from 0 to a_d in a
        # matter + radiation universe with R proportional to a
        aeq = 1.0 / p["zeq"]
        ad = aeq * p["Rd"] / p["Req"]
        H0_over_c = p["keq"] * np.sqrt(aeq / 2 / c.Omega_m)
        val, _ = quad(lambda a: 1.0 / (np.sqrt(3 * (1 + p["Req"] * a / aeq)) * np.sqrt(a + aeq)), 0.0, ad,
                      epsabs=0, epsrel=1e-13, limit=200)
        s_int = val / (H0_over_c * np.sqrt(c.Omega_m))
        assert abs(p["s"] / s_int - 1) < 1e-9, ("s integral", p["s"], s_int)
        # eq. 26 fits eq. 6 to a few per cent for these parameters
        s_fit = 44.5 * np.log(9.83 / wm) / np.sqrt(1 + 10 * wb ** 0.75)
        assert abs(p["s"] / s_fit - 1) < 0.025, ("s fit", p["s"], s_fit)
        # eq. 21: the baryon transfer function vanishes where k s~(k) = m pi
        kk = np.logspace(-3, 0.5, 4000)
        pk = parts_fn(c, kk)
        g = pk["kMpc"] * pk["s_tilde"]
        for m in range(1, 7):
            i = np.nonzero((g[:-1] < m * np.pi) & (g[1:] >= m * np.pi))[0][0]
            km = brentq(lambda k: float(parts_fn(c, np.array([k]))["kMpc"][0] *
                                        parts_fn(c, np.array([k]))["s_tilde"][0]) - m * np.pi, kk[i], kk[i + 1],
                        xtol=1e-15)
            tb = parts_fn(c, np.array([km * (1 - 1e-3), km, km * (1 + 1e-3)]))["Tb"]
            assert abs(tb[1]) < 1e-9 * np.abs(pk["Tb"]).max(), ("node", m, tb)
            assert tb[0] * tb[2] < 0, ("sign change", m, tb)
        # no-wiggle form (eqs. 26-31): the same smooth shape
        k = np.logspace(-4, 1.5, 300)
        s = s_fit
        r = parts_fn(c, k)["T"] / _no_wiggle(c, k)
        damped = k > 0.3
        assert np.abs(r[damped] - 1).max() < 0.02, ("no-wiggle, damped", np.abs(r[damped] - 1).max())
        assert np.abs(r[~damped] - 1).max() < 0.05, ("no-wiggle, oscillating", np.abs(r[~damped] - 1).max())
        # averaged over one acoustic period 2 pi / s the wiggles cancel and
        # the full form follows the no-wiggle one closely
        k = np.linspace(0.005, 0.6, 6000)
        r = parts_fn(c, k)["T"] / _no_wiggle(c, k)
        w = int(round(2 * np.pi / s / c.h / (k[1] - k[0])))
        r_avg = np.convolve(r, np.ones(w) / w, mode="valid")
        assert np.abs(r_avg - 1).max() < 0.0145, ("period average", np.abs(r_avg - 1).max())
    # zero-baryon limit: eq. 16 -> T0(k, 1, 1) (eqs. 10, 19, 20)
    c0 = oc.Cosmology(Omega_c=0.3, Omega_b=1e-13, h=0.7)
    k = np.logspace(-4, 2, 200)
    p0 = parts_fn(c0, k)
    q = k * c0.h / (13.41 * p0["keq"])
    L = np.log(np.e + 1.8 * q)
    T0 = L / (L + (14.2 + 386 / (1 + 69.9 * q ** 1.08)) * q * q)
    assert np.abs(p0["T"] / T0 - 1).max() < 1e-10, ("zero baryon", np.abs(p0["T"] / T0 - 1).max())
    if lmp_fn is None:
        return
    for c in COSMOS[:2]:
        # sigma8 from an independent quadrature (trapezoid in ln k, not the
        # oracle's adaptive quad) of the normalised P(k) at a = 1
        lnk = np.linspace(np.log(1e-5), np.log(2e2), 400001)
        kk = np.exp(lnk)
        x = kk * 8.0
        W = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        s2 = np.trapezoid(kk ** 3 * lmp_fn(c, kk) * W * W, lnk) / (2 * np.pi ** 2)
        assert abs(s2 / c.sigma8 ** 2 - 1) < 1e-6, ("sigma8", np.sqrt(s2), c.sigma8)
        # primordial slope where T = 1
        k1, k2 = 1e-7, 2e-7
        slope = np.log(lmp_fn(c, k2) / lmp_fn(c, k1)) / np.log(k2 / k1)
        assert abs(slope - c.n_s) < 1e-5, ("slope", slope)
        # P(k, a) = D(a)^2 P(k, 1)
        kk = np.logspace(-3, 1, 20)
        for a in (0.1, 0.5):
            np.testing.assert_allclose(lmp_fn(c, kk, a), lmp_fn(c, kk) * oc.growth_factor(c, a) ** 2, rtol=1e-12)


def test_eh98_pins():
    _pins(oc.eh98_parts, oc.linear_matter_power)


# Plausible typos: a digit swapped or dropped in a coefficient, a sign or
# exponent flipped, the wrong wave number in the node structure.
MUTATIONS = [
    ("7.46e-2 * wm", "7.64e-2 * wm"),
    ("2.50e4 * wm", "2.05e4 * wm"),
    ("1291 * wm", "1219 * wm"),
    ("Rd = 31.5 * wb", "Rd = 35.1 * wb"),
    ("Req = 31.5 * wb", "Req = 35.1 * wb"),
    ("np.sqrt(6 / Req)", "np.sqrt(6 * Req)"),
    ("(1 + np.sqrt(Req))", "(1 - np.sqrt(Req))"),
    ("2.07 * keq", "2.70 * keq"),
    ("(1 + Rd) ** -0.75", "(1 + Rd) ** 0.75"),
    ("(beta_node / (k * s)) ** 3", "(beta_node / (k * s)) ** 2"),
    ("1.6 * wb ** 0.52", "6.1 * wb ** 0.52"),
    ("(k / ksilk) ** 1.4", "(k / ksilk) ** 4.1"),
    ("(k * s / 5.2) ** 2", "(k * s / 2.5) ** 2"),
    ("(k * s / 5.4) ** 4", "(k * s / 4.5) ** 4"),
    ("14.2 / alpha", "12.4 / alpha"),
    ("386 / (1 + 69.9", "368 / (1 + 69.9"),
    ("13.41 * keq", "14.31 * keq"),
    ("a2 ** -(fb ** 3)", "a2 ** -(fb ** 2)"),
    ("0.944 / (1 + (458 * wm)", "0.494 / (1 + (458 * wm)"),
    ("(0.395 * wm) ** -0.0266", "(0.395 * wm) ** -0.266"),
    ("G = y * (-6 * sq", "G = y * (-3 * sq"),
    ("1.8 * beta * q", "8.1 * beta * q"),
    ("fb * tb + fc * tc", "fc * tb + fb * tc"),
    ("k = np.asarray(k, np.float64) * c.h", "k = np.asarray(k, np.float64) / c.h"),
]

# Typos that move T(k) by less than the pins above resolve (the largest
# relative change of T over k in [1e-3, 10] h/Mpc for the cosmologies
# above), whose coefficients therefore rest on the transcription alone:
#   z_d's 0.659 -> 0.569 (eq. 4): 0.37%;  beta_node's 8.41 -> 8.14 (eq. 23):
#   0.29%;  beta_b's 17.2 -> 12.7 (eq. 24): 0.83%;  beta_b's (3 - 2 f_b) ->
#   (3 + 2 f_b): 0.78%;  alpha_c's 46.9 -> 49.6 (eq. 11): 0.62%.

LMP_MUTATIONS = [
    ("norm * k ** c.n_s", "norm * k ** (c.n_s - 1)"),
    ("_sigma2_unnorm(c, 8.0)", "_sigma2_unnorm(c, 8.0 / c.h)"),
    ("return val / (2 * np.pi ** 2)", "return val / (2 * np.pi)"),
    ("k ** 3 * k ** c.n_s", "k ** 2 * k ** c.n_s"),
    ("growth_factor(c, a) ** 2", "growth_factor(c, a)"),
]


def _mutant(fn_names, old, new):
    src = "\n".join(textwrap.dedent(inspect.getsource(getattr(oc, f))) for f in fn_names)
    assert src.count(old) == 1, (old, src.count(old))
    ns = dict(vars(oc))
    exec(compile(src.replace(old, new), "<mutant>", "exec"), ns)
    return ns


@pytest.mark.parametrize("old,new", MUTATIONS)
def test_pins_catch_eh98_typo(old, new):
    ns = _mutant(["eh98_parts"], old, new)
    with pytest.raises(AssertionError):
        _pins(ns["eh98_parts"])


@pytest.mark.parametrize("old,new", LMP_MUTATIONS)
def test_pins_catch_power_spectrum_typo(old, new):
    ns = _mutant(["eh98_parts", "eisenstein_hu", "_sigma2_unnorm", "linear_matter_power"], old, new)
    with pytest.raises(AssertionError):
        _pins(ns["eh98_parts"], ns["linear_matter_power"])
