"""Host-side background cosmology, growth and linear power (float64 NumPy).

JaxPM takes these from jax_cosmo (``jaxpm/growth.py:6-9``, ``pm.py:2,78``,
``ode.py:2,140,143``; tests ``conftest.py:56-70,90-91``).  jax_cosmo is not
part of the reference tree, so this module restates its published algorithm:

* ``Esqr(a) = Om a^-3 + Ok a^-2 + Ode exp(f_de(a))``,
  ``f_de(a) = -3 (1 + w0 + wa) ln a + 3 wa (a - 1)`` (CPL dark energy).
* Growth: ``D'' = -q D' + r D`` and ``D2'' = -q D2' + r D2 - r D^2`` with
  ``q = (2 - (Om(a) + (1 + 3 w(a)) Ode(a)) / 2) / a`` and ``r = 1.5 Om(a)/a^2``,
  from ``D = a0, D' = 1, D2 = -3/7 a0^2, D2' = -6/7 a0`` at ``a0 = 1e-3``,
  tabulated on ``logspace(-3, 0, 256)`` and normalised to 1 at ``a = 1``
  (so ``-D2_true = 3/7 growth_factor_second``, pm.py:117).  Values between
  nodes are linear interpolations in ``a``; ``growth_factor`` is clipped to
  [0, 1].  The ODE is integrated here with classical RK4 in ``ln a``, 32
  substeps per table interval.
* Eisenstein & Hu (1998) transfer function with baryon wiggles
  (``T_cmb = 2.7255``), primordial slope ``n_s``, normalised to ``sigma8``
  with a real-space top hat of radius 8 Mpc/h.

These are scalar, host-side coefficients (one value per time step); no GPU
kernel is involved.
"""
from dataclasses import dataclass, field

import numpy as np

T_CMB = 2.7255
_LOG10_AMIN = -3.0
_STEPS = 256
_SUBSTEPS = 32


@dataclass(frozen=True)
class Cosmology:
    """jax_cosmo.Cosmology's parameters (Planck15 defaults)."""
    Omega_c: float = 0.2589
    Omega_b: float = 0.0486
    h: float = 0.6774
    n_s: float = 0.9667
    sigma8: float = 0.8159
    Omega_k: float = 0.0
    w0: float = -1.0
    wa: float = 0.0
    _cache: dict = field(default_factory=dict, compare=False, hash=False, repr=False)

    @property
    def Omega_m(self):
        return self.Omega_c + self.Omega_b

    @property
    def Omega_de(self):
        return 1.0 - self.Omega_k - self.Omega_m

    def replace(self, **kw):
        d = {k: getattr(self, k) for k in ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "Omega_k", "w0", "wa")}
        d.update(kw)
        return Cosmology(**d)


def Planck15(**kw):
    return Cosmology(**kw)


def Planck18(**kw):
    """The parameter set of the reference's tests (conftest.py:57-68)."""
    base = dict(Omega_c=0.2607, Omega_b=0.0490, Omega_k=0.0, h=0.6766, n_s=0.9665, sigma8=0.8102, w0=-1.0,
                wa=0.0)
    base.update(kw)
    return Cosmology(**base)


# -------------------------------------------------------------- background

def w(cosmo, a):
    return cosmo.w0 + (1.0 - np.asarray(a, np.float64)) * cosmo.wa


def f_de(cosmo, a):
    a = np.asarray(a, np.float64)
    return -3.0 * (1.0 + cosmo.w0 + cosmo.wa) * np.log(a) + 3.0 * cosmo.wa * (a - 1.0)


def Esqr(cosmo, a):
    a = np.asarray(a, np.float64)
    return cosmo.Omega_m * a ** -3 + cosmo.Omega_k * a ** -2 + cosmo.Omega_de * np.exp(f_de(cosmo, a))


def Omega_m_a(cosmo, a):
    return cosmo.Omega_m * np.asarray(a, np.float64) ** -3 / Esqr(cosmo, a)


def Omega_de_a(cosmo, a):
    return cosmo.Omega_de * np.exp(f_de(cosmo, a)) / Esqr(cosmo, a)


# ------------------------------------------------------------------ growth

def _qr(cosmo, a):
    """Coefficients of the growth ODE at scale factor(s) a:
    d2D/da2 = -q dD/da + r D (+ the 2LPT source -r D1^2 for D2)."""
    om = Omega_m_a(cosmo, a)
    q = (2.0 - 0.5 * (om + (1.0 + 3.0 * w(cosmo, a)) * Omega_de_a(cosmo, a))) / a
    r = 1.5 * om / (a * a)
    return q, r


def growth_tables(cosmo):
    """(atab, gtab, ftab, htab, g2tab, f2tab, h2tab) -- the 7-tuple that
    growth.py:218-221 and 255-258 index.

    Classical RK4 in ln a on 256 log-spaced nodes from a = 1e-3 with 32 fixed
    substeps per interval, for y = (D1, D2, dD1/da, dD2/da) with growing-mode
    initial values (D1 = a, D2 = -3/7 a^2).  The ODE coefficients are
    evaluated for every RK4 stage at once (vectorised over the fixed step
    grid); the stages themselves run on scalars.  A fixed step makes the
    tables smooth in the cosmological parameters."""
    t = cosmo._cache.get("growth")
    if t is not None:
        return t
    atab = np.logspace(_LOG10_AMIN, 0.0, _STEPS)
    lna = np.log(atab)
    # ln a of every substep start, by the same repeated addition as the steps
    xs = np.empty((_STEPS - 1, _SUBSTEPS))
    hs = np.empty(_STEPS - 1)
    for i in range(_STEPS - 1):
        h = (lna[i + 1] - lna[i]) / _SUBSTEPS
        hs[i] = h
        x = lna[i]
        for j in range(_SUBSTEPS):
            xs[i, j] = x
            x = x + h
    hh = hs[:, None]
    a0, am, a1 = np.exp(xs), np.exp(xs + 0.5 * hh), np.exp(xs + hh)
    q0, r0 = _qr(cosmo, a0)
    qm, rm = _qr(cosmo, am)
    q1, r1 = _qr(cosmo, a1)
    qn, rn = _qr(cosmo, atab)
    a0, am, a1 = a0.tolist(), am.tolist(), a1.tolist()
    q0, r0, qm, rm, q1, r1 = q0.tolist(), r0.tolist(), qm.tolist(), rm.tolist(), q1.tolist(), r1.tolist()

    def f(A, Q, R, d1, d2, v1, v2):
        acc1 = -Q * v1 + R * d1
        acc2 = -Q * v2 + R * d2 - R * d1 * d1
        return v1 * A, v2 * A, acc1 * A, acc2 * A

    def accel(Q, R, d1, d2, v1, v2):
        return -Q * v1 + R * d1, -Q * v2 + R * d2 - R * d1 * d1

    d1, d2, v1, v2 = float(atab[0]), -3.0 / 7.0 * float(atab[0]) ** 2, 1.0, -6.0 / 7.0 * float(atab[0])
    ys = np.empty((_STEPS, 4))
    acc = np.empty((_STEPS, 2))
    ys[0] = (d1, d2, v1, v2)
    acc[0] = accel(float(qn[0]), float(rn[0]), d1, d2, v1, v2)
    for i in range(_STEPS - 1):
        h = float(hs[i])
        hh2 = 0.5 * h
        h6 = h / 6.0
        A0, AM, A1 = a0[i], am[i], a1[i]
        Q0, R0, QM, RM, Q1, R1 = q0[i], r0[i], qm[i], rm[i], q1[i], r1[i]
        for j in range(_SUBSTEPS):
            k1 = f(A0[j], Q0[j], R0[j], d1, d2, v1, v2)
            k2 = f(AM[j], QM[j], RM[j], d1 + hh2 * k1[0], d2 + hh2 * k1[1], v1 + hh2 * k1[2], v2 + hh2 * k1[3])
            k3 = f(AM[j], QM[j], RM[j], d1 + hh2 * k2[0], d2 + hh2 * k2[1], v1 + hh2 * k2[2], v2 + hh2 * k2[3])
            k4 = f(A1[j], Q1[j], R1[j], d1 + h * k3[0], d2 + h * k3[1], v1 + h * k3[2], v2 + h * k3[3])
            d1 = d1 + h6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            d2 = d2 + h6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
            v1 = v1 + h6 * (k1[2] + 2 * k2[2] + 2 * k3[2] + k4[2])
            v2 = v2 + h6 * (k1[3] + 2 * k2[3] + 2 * k3[3] + k4[3])
        ys[i + 1] = (d1, d2, v1, v2)
        acc[i + 1] = accel(float(qn[i + 1]), float(rn[i + 1]), d1, d2, v1, v2)
    d1, d2, v1, v2 = ys.T
    gtab = d1 / d1[-1]
    g2tab = d2 / d2[-1]
    ftab = v1 / d1[-1] * atab / gtab
    f2tab = v2 / d2[-1] * atab / g2tab
    htab = acc[:, 0] / d1[-1] * atab / gtab
    h2tab = acc[:, 1] / d2[-1] * atab / g2tab
    t = (atab, gtab, ftab, htab, g2tab, f2tab, h2tab)
    cosmo._cache["growth"] = t
    return t


def growth_factor(cosmo, a):
    t = growth_tables(cosmo)
    return np.clip(np.interp(a, t[0], t[1]), 0.0, 1.0)


def growth_rate(cosmo, a):
    t = growth_tables(cosmo)
    return np.interp(a, t[0], t[2])


def growth_factor_second(cosmo, a):
    t = growth_tables(cosmo)
    return np.interp(a, t[0], t[4])


def growth_rate_second(cosmo, a):
    t = growth_tables(cosmo)
    return np.interp(a, t[0], t[5])


# --------------------------------------------------------------- P(k), EH98

def eisenstein_hu(cosmo, k):
    """EH98 transfer function (with baryon oscillations); ``k`` in h/Mpc."""
    k = np.asarray(k, np.float64) * cosmo.h
    theta = T_CMB / 2.7
    om_h2 = cosmo.Omega_m * cosmo.h ** 2
    ob_h2 = cosmo.Omega_b * cosmo.h ** 2
    f_b = cosmo.Omega_b / cosmo.Omega_m
    f_c = cosmo.Omega_c / cosmo.Omega_m
    z_eq = 2.50e4 * om_h2 * theta ** -4                                  # eq. 2
    k_eq = 7.46e-2 * om_h2 * theta ** -2                                 # eq. 3
    b1 = 0.313 * om_h2 ** -0.419 * (1.0 + 0.607 * om_h2 ** 0.674)
    b2 = 0.238 * om_h2 ** 0.223
    z_d = 1291.0 * om_h2 ** 0.251 / (1.0 + 0.659 * om_h2 ** 0.828) * (1.0 + b1 * ob_h2 ** b2)  # eq. 4
    r_of = lambda z: 31.5 * ob_h2 * theta ** -4 * (1e3 / z)             # eq. 5
    R_d, R_eq = r_of(z_d), r_of(z_eq)
    s = (2.0 / (3.0 * k_eq)) * np.sqrt(6.0 / R_eq) * np.log(
        (np.sqrt(1.0 + R_d) + np.sqrt(R_d + R_eq)) / (1.0 + np.sqrt(R_eq)))  # eq. 6
    k_silk = 1.6 * ob_h2 ** 0.52 * om_h2 ** 0.73 * (1.0 + (10.4 * om_h2) ** -0.95)  # eq. 7
    q = k / (13.41 * k_eq)                                               # eq. 10
    alpha_c = ((46.9 * om_h2) ** 0.670 * (1.0 + (32.1 * om_h2) ** -0.532)) ** -f_b * \
        ((12.0 * om_h2) ** 0.424 * (1.0 + (45.0 * om_h2) ** -0.582)) ** -(f_b ** 3)  # eq. 11
    beta_c = 1.0 / (1.0 + 0.944 / (1.0 + (458.0 * om_h2) ** -0.708) *
                    (f_c ** ((0.395 * om_h2) ** -0.0266) - 1.0))          # eq. 12

    def t_tilde(alpha, beta):                                            # eqs. 19-20
        L = np.log(np.e + 1.8 * beta * q)
        C = 14.2 / alpha + 386.0 / (1.0 + 69.9 * q ** 1.08)
        return L / (L + C * q ** 2)

    f = 1.0 / (1.0 + (k * s / 5.4) ** 4)                                 # eq. 18
    T_c = f * t_tilde(1.0, beta_c) + (1.0 - f) * t_tilde(alpha_c, beta_c)  # eq. 17
    y = (1.0 + z_eq) / (1.0 + z_d)
    G = y * (-6.0 * np.sqrt(1.0 + y) + (2.0 + 3.0 * y) *
             np.log((np.sqrt(1.0 + y) + 1.0) / (np.sqrt(1.0 + y) - 1.0)))  # eq. 15
    alpha_b = 2.07 * k_eq * s * (1.0 + R_d) ** -0.75 * G                 # eq. 14
    beta_node = 8.41 * om_h2 ** 0.435                                    # eq. 23
    s_tilde = s / (1.0 + (beta_node / (k * s)) ** 3) ** (1.0 / 3.0)      # eq. 22
    beta_b = 0.5 + f_b + (3.0 - 2.0 * f_b) * np.sqrt((17.2 * om_h2) ** 2 + 1.0)  # eq. 24
    x = k * s_tilde
    with np.errstate(invalid="ignore", divide="ignore"):
        j0 = np.where(x > 0, np.sin(x) / np.where(x > 0, x, 1.0), 1.0)
    T_b = (t_tilde(1.0, 1.0) / (1.0 + (k * s / 5.2) ** 2) +
           alpha_b / (1.0 + (beta_b / (k * s)) ** 3) * np.exp(-(k / k_silk) ** 1.4)) * j0  # eq. 21
    return f_b * T_b + f_c * T_c                                         # eq. 16


def _sigma8_unnormalised(cosmo):
    v = cosmo._cache.get("sigma8_raw")
    if v is not None:
        return v
    # Composite Simpson in ln k over [1e-5, 1e3] h/Mpc, 2^14 intervals.
    lnk = np.linspace(np.log(1e-5), np.log(1e3), 2 ** 14 + 1)
    k = np.exp(lnk)
    x = 8.0 * k
    wt = 3.0 * (np.sin(x) - x * np.cos(x)) / x ** 3
    f = k ** (3.0 + cosmo.n_s) * eisenstein_hu(cosmo, k) ** 2 * wt ** 2
    h = lnk[1] - lnk[0]
    integral = h / 3.0 * (f[0] + f[-1] + 4.0 * f[1:-1:2].sum() + 2.0 * f[2:-1:2].sum())
    v = integral / (2.0 * np.pi ** 2)
    cosmo._cache["sigma8_raw"] = v
    return v


def linear_matter_power(cosmo, k, a=1.0):
    """sigma8-normalised linear matter power [(Mpc/h)^3] at scale factor a."""
    k = np.asarray(k, np.float64)
    norm = cosmo.sigma8 ** 2 / _sigma8_unnormalised(cosmo)
    return norm * k ** cosmo.n_s * eisenstein_hu(cosmo, k) ** 2 * growth_factor(cosmo, a) ** 2


def pk_table(cosmo, kmin=1e-4, kmax=1e1, n=128):
    """The (k, P(k)) table the reference's tests interpolate linearly
    (``jnp.logspace(-4, 1, 128)``, conftest.py:90-94)."""
    k = np.logspace(np.log10(kmin), np.log10(kmax), n)
    return k, linear_matter_power(cosmo, k)
