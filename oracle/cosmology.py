"""Oracle background cosmology and growth (test infrastructure only).

JaxPM takes these from jax_cosmo (``growth.py:6-9``; ``pm.py:2,78``;
``ode.py:2,140,143``), which is not vendored in the reference and is not
installed here.  This module restates jax_cosmo's published algorithm in
float64:

* ``Esqr(a) = Om a^-3 + Ok a^-2 + Ode exp(f_de(a))`` with the CPL dark-energy
  exponent ``f_de(a) = -3 (1 + w0 + wa) ln a + 3 wa (a - 1)``.
* First and second order growth from the ODEs
  ``D'' = -q D' + r D`` and ``D2'' = -q D2' + r D2 - r D^2`` with
  ``q = (2 - (Om(a) + (1 + 3w) Ode(a)) / 2) / a``, ``r = 1.5 Om(a) / a^2``,
  started at ``a0 = 1e-3`` from ``D = a0, D' = 1, D2 = -3/7 a0^2,
  D2' = -6/7 a0``, tabulated on ``logspace(-3, 0, 256)``, normalised to 1 at
  a = 1 (so ``-D2_true = 3/7 D2``, pm.py:117) and linearly interpolated in
  ``a``; ``growth_factor`` is clipped to [0, 1].
* The growth-table derivatives used by ``growth.py`` (``htab``, ``h2tab`` =
  a D''/D) interpolated in ``log a`` as growth.py:218-221, 255-258 do.
* Eisenstein & Hu (1998, ApJ 496, 605) transfer function with baryon
  oscillations (their eqs. 2-24), ``T_cmb = 2.7255``, primordial ``k^n_s``,
  normalised to ``sigma8`` with a real-space top-hat of 8 Mpc/h.

The oracle integrates the ODE with ``scipy.integrate.solve_ivp`` (DOP853,
rtol 1e-12) at the table nodes.

Parity pinned: Einstein-de Sitter limits and the Heath integral
``D1 ~ H(a) int_0^a da' / (a' H(a'))^3`` for flat LCDM; EH98 by CODATA
constants (k_eq, z_eq, R), the sound-horizon integral and the paper's fit,
the zero-baryon limit, the nodes of T_b, the no-wiggle form, an independent
sigma8 quadrature and a mutation test of planted typos
(tests/test_oracle_eh98.py).
Parity unpinned: the exact values of jax_cosmo's tables (its ODE tolerance,
table length and Romberg sigma8 quadrature) -- jax_cosmo cannot be run here;
five EH98 coefficients that move T(k) by less than 1% (listed in that test).
"""
from dataclasses import dataclass

import numpy as np
from scipy.integrate import quad, solve_ivp

T_CMB = 2.7255
A_MIN_LOG10 = -3.0
TABLE_STEPS = 256


@dataclass(frozen=True)
class Cosmology:
    Omega_c: float = 0.2589
    Omega_b: float = 0.0486
    h: float = 0.6774
    n_s: float = 0.9667
    sigma8: float = 0.8159
    Omega_k: float = 0.0
    w0: float = -1.0
    wa: float = 0.0

    @property
    def Omega_m(self):
        return self.Omega_c + self.Omega_b

    @property
    def Omega_de(self):
        return 1.0 - self.Omega_k - self.Omega_m


def f_de(c, a):
    return -3.0 * (1.0 + c.w0 + c.wa) * np.log(a) + 3.0 * c.wa * (a - 1.0)


def w(c, a):
    return c.w0 + (1.0 - a) * c.wa


def Esqr(c, a):
    a = np.asarray(a, np.float64)
    return c.Omega_m * a ** -3 + c.Omega_k * a ** -2 + c.Omega_de * np.exp(f_de(c, a))


def E(c, a):
    """growth.py:31-52."""
    return np.sqrt(Esqr(c, a))


def Omega_m_a(c, a):
    return c.Omega_m * np.asarray(a, np.float64) ** -3 / Esqr(c, a)


def Omega_de_a(c, a):
    return c.Omega_de * np.exp(f_de(c, a)) / Esqr(c, a)


def _rhs(c):
    def f(a, y):
        d1, d2, v1, v2 = y
        q = (2.0 - 0.5 * (Omega_m_a(c, a) + (1.0 + 3.0 * w(c, a)) * Omega_de_a(c, a))) / a
        r = 1.5 * Omega_m_a(c, a) / a / a
        return [v1, v2, -q * v1 + r * d1, -q * v2 + r * d2 - r * d1 * d1]
    return f


_TABLES = {}


def growth_tables(c):
    """(atab, gtab, ftab, htab, g2tab, f2tab, h2tab), the 7-tuple layout of
    growth.py:219 / 256."""
    if c in _TABLES:
        return _TABLES[c]
    atab = np.logspace(A_MIN_LOG10, 0.0, TABLE_STEPS)
    a0 = atab[0]
    y0 = [a0, -3.0 / 7.0 * a0 ** 2, 1.0, -6.0 / 7.0 * a0]
    sol = solve_ivp(_rhs(c), (a0, 1.0), y0, t_eval=atab, method='DOP853',
                    rtol=1e-12, atol=1e-15)
    d1, d2, v1, v2 = sol.y
    acc = np.array([_rhs(c)(a, y)[2:] for a, y in zip(atab, sol.y.T)])
    gtab = d1 / d1[-1]
    g2tab = d2 / d2[-1]
    ftab = v1 / d1[-1] * atab / gtab
    f2tab = v2 / d2[-1] * atab / g2tab
    htab = acc[:, 0] / d1[-1] * atab / gtab
    h2tab = acc[:, 1] / d2[-1] * atab / g2tab
    t = (atab, gtab, ftab, htab, g2tab, f2tab, h2tab)
    _TABLES[c] = t
    return t


def growth_factor(c, a):
    t = growth_tables(c)
    return np.clip(np.interp(a, t[0], t[1]), 0.0, 1.0)


def growth_rate(c, a):
    t = growth_tables(c)
    return np.interp(a, t[0], t[2])


def growth_factor_second(c, a):
    t = growth_tables(c)
    return np.interp(a, t[0], t[4])


def growth_rate_second(c, a):
    t = growth_tables(c)
    return np.interp(a, t[0], t[5])


def df_de(c, a, epsilon=1e-5):
    """growth.py:55-85."""
    return (3 * c.wa * (np.log(a - epsilon) - (a - 1) / (a - epsilon)) /
            np.log(a - epsilon) ** 2)


def dEa(c, a):
    """growth.py:88-116."""
    return (0.5 * (-3 * c.Omega_m * a ** -4.0 - 2 * c.Omega_k * a ** -3.0 +
                   df_de(c, a) * c.Omega_de * np.exp(f_de(c, a))) / E(c, a))


def gp(c, a):
    """growth.py:264-293: dD1/da = f1 D1 / a."""
    return growth_rate(c, a) * growth_factor(c, a) / a


def Gf(c, a):
    """growth.py:124-154."""
    return gp(c, a) * a ** 3 * E(c, a)


def Gf2(c, a):
    """growth.py:157-187."""
    return growth_rate_second(c, a) * growth_factor_second(c, a) / a * a ** 3 * E(c, a)


def dGfa(c, a):
    """growth.py:190-224 (f1p interpolated in log a, 221)."""
    t = growth_tables(c)
    d1f = growth_rate(c, a) * growth_factor(c, a) / a
    f1p = np.interp(np.log(a), np.log(t[0]), t[3] / t[0] * t[1])
    ea = E(c, a)
    return f1p * a ** 3 * ea + d1f * a ** 3 * dEa(c, a) + 3 * a ** 2 * ea * d1f


def dGf2a(c, a):
    """growth.py:227-261 (f2p interpolated in log a, 258)."""
    t = growth_tables(c)
    d2f = growth_rate_second(c, a) * growth_factor_second(c, a) / a
    f2p = np.interp(np.log(a), np.log(t[0]), t[6] / t[0] * t[4])
    ea = E(c, a)
    return f2p * a ** 3 * ea + d2f * a ** 3 * dEa(c, a) + 3 * a ** 2 * ea * d2f


# ---------------------------------------------------------------- EH98 P(k)

def eh98_parts(c, k):
    """EH98 transfer function and its intermediate quantities; ``k`` in h/Mpc.

    Returns a dict: ``T`` (eq. 16), ``Tb`` (eq. 21), ``Tc`` (eq. 17), the
    sound horizon ``s`` (eq. 6, Mpc), ``s_tilde`` (eq. 22), ``keq`` (eq. 3,
    1/Mpc), ``zeq``, ``zd``, ``Rd``, ``Req``, ``ksilk``, ``alpha_c``,
    ``beta_c``, ``alpha_b``, ``beta_b``, ``beta_node`` and ``kMpc`` (k in
    1/Mpc), so each equation can be pinned on its own.
    """
    k = np.asarray(k, np.float64) * c.h  # 1/Mpc
    th = T_CMB / 2.7
    wm = c.Omega_m * c.h ** 2
    wb = c.Omega_b * c.h ** 2
    fb = c.Omega_b / c.Omega_m
    fc = c.Omega_c / c.Omega_m
    zeq = 2.50e4 * wm * th ** -4                                   # eq. 2
    keq = 7.46e-2 * wm * th ** -2                                  # eq. 3
    b1 = 0.313 * wm ** -0.419 * (1 + 0.607 * wm ** 0.674)          # eq. 4
    b2 = 0.238 * wm ** 0.223
    zd = 1291 * wm ** 0.251 / (1 + 0.659 * wm ** 0.828) * (1 + b1 * wb ** b2)
    Rd = 31.5 * wb * th ** -4 * (zd / 1e3) ** -1                   # eq. 5
    Req = 31.5 * wb * th ** -4 * (zeq / 1e3) ** -1
    s = 2 / (3 * keq) * np.sqrt(6 / Req) * np.log(                 # eq. 6
        (np.sqrt(1 + Rd) + np.sqrt(Rd + Req)) / (1 + np.sqrt(Req)))
    ksilk = 1.6 * wb ** 0.52 * wm ** 0.73 * (1 + (10.4 * wm) ** -0.95)   # eq. 7
    q = k / (13.41 * keq)                                          # eq. 10
    a1 = (46.9 * wm) ** 0.670 * (1 + (32.1 * wm) ** -0.532)        # eq. 11
    a2 = (12.0 * wm) ** 0.424 * (1 + (45.0 * wm) ** -0.582)
    alpha_c = a1 ** -fb * a2 ** -(fb ** 3)
    bb1 = 0.944 / (1 + (458 * wm) ** -0.708)                       # eq. 12
    bb2 = (0.395 * wm) ** -0.0266
    beta_c = 1 / (1 + bb1 * (fc ** bb2 - 1))

    def t0(alpha, beta):                                           # eqs. 19-20
        lg = np.log(np.e + 1.8 * beta * q)
        cc = 14.2 / alpha + 386 / (1 + 69.9 * q ** 1.08)
        return lg / (lg + cc * q * q)

    f = 1 / (1 + (k * s / 5.4) ** 4)                               # eq. 18
    tc = f * t0(1.0, beta_c) + (1 - f) * t0(alpha_c, beta_c)       # eq. 17
    y = (1 + zeq) / (1 + zd)
    sq = np.sqrt(1 + y)
    G = y * (-6 * sq + (2 + 3 * y) * np.log((sq + 1) / (sq - 1)))  # eq. 15
    alpha_b = 2.07 * keq * s * (1 + Rd) ** -0.75 * G               # eq. 14
    beta_node = 8.41 * wm ** 0.435                                 # eq. 23
    st = s / (1 + (beta_node / (k * s)) ** 3) ** (1 / 3)           # eq. 22
    beta_b = 0.5 + fb + (3 - 2 * fb) * np.sqrt((17.2 * wm) ** 2 + 1)    # eq. 24
    ks = k * st
    j0 = np.where(ks == 0, 1.0, np.sin(ks) / np.where(ks == 0, 1, ks))
    tb = (t0(1.0, 1.0) / (1 + (k * s / 5.2) ** 2) +                # eq. 21
          alpha_b / (1 + (beta_b / (k * s)) ** 3) * np.exp(-(k / ksilk) ** 1.4)) * j0
    return dict(T=fb * tb + fc * tc, Tb=tb, Tc=tc, s=s, s_tilde=st, keq=keq, zeq=zeq, zd=zd, Rd=Rd, Req=Req,
                ksilk=ksilk, alpha_c=alpha_c, beta_c=beta_c, alpha_b=alpha_b, beta_b=beta_b,
                beta_node=beta_node, kMpc=k)


def eisenstein_hu(c, k):
    """EH98 transfer function with baryon oscillations (eq. 16); ``k`` in h/Mpc."""
    return eh98_parts(c, k)["T"]


def _sigma2_unnorm(c, R):
    def integrand(lnk):
        k = np.exp(lnk)
        x = k * R
        wt = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        return k ** 3 * k ** c.n_s * eisenstein_hu(c, k) ** 2 * wt * wt
    val, _ = quad(integrand, np.log(1e-5), np.log(1e3), limit=2000, epsabs=0, epsrel=1e-11)
    return val / (2 * np.pi ** 2)


def linear_matter_power(c, k, a=1.0):
    """sigma8-normalised linear P(k) [(Mpc/h)^3] at scale factor ``a``."""
    k = np.asarray(k, np.float64)
    norm = c.sigma8 ** 2 / _sigma2_unnorm(c, 8.0)
    return norm * k ** c.n_s * eisenstein_hu(c, k) ** 2 * growth_factor(c, a) ** 2


def heath_growth(c, a):
    """Unnormalised LCDM growing mode H(a) int_0^a da'/(a' H)^3 (flat, w=-1)."""
    def integrand(x):
        return 1.0 / (x * E(c, x)) ** 3
    val, _ = quad(integrand, 0.0, a, epsabs=0, epsrel=1e-13, limit=500)
    return 2.5 * c.Omega_m * E(c, a) * val
