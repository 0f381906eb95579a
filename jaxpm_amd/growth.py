"""Scalar time coefficients of the PM step, mirroring ``jaxpm/growth.py``.

Host-side float64 NumPy over the growth tables of
:mod:`jaxpm_amd.cosmology` (the jax_cosmo functions that jaxpm re-exports,
growth.py:6-9).  Formulas follow growth.py line by line (cited per function).
"""
import numpy as np

from .cosmology import (Esqr, Omega_de_a, Omega_m_a, f_de, growth_factor,  # noqa: F401
                        growth_factor_second, growth_rate, growth_rate_second,
                        growth_tables, w)

__all__ = [
    "growth_factor", "growth_rate", "growth_factor_second", "growth_rate_second",
    "E", "df_de", "dEa", "Gf", "Gf2", "dGfa", "dGf2a", "gp", "Dplusdada", "Dplus_to_a",
]


def E(cosmo, a):
    """growth.py:31-52: sqrt(Esqr)."""
    return np.sqrt(Esqr(cosmo, a))


def df_de(cosmo, a, epsilon=1e-5):
    """growth.py:55-85."""
    a = np.asarray(a, np.float64)
    return (3 * cosmo.wa * (np.log(a - epsilon) - (a - 1) / (a - epsilon)) /
            np.power(np.log(a - epsilon), 2))


def dEa(cosmo, a):
    """growth.py:88-116."""
    a = np.asarray(a, np.float64)
    return (0.5 * (-3 * cosmo.Omega_m * np.power(a, -4) - 2 * cosmo.Omega_k * np.power(a, -3) +
                   df_de(cosmo, a) * cosmo.Omega_de * np.exp(f_de(cosmo, a))) / E(cosmo, a))


def gp(cosmo, a):
    """growth.py:264-293: dD1/da = f1 D1 / a."""
    return growth_rate(cosmo, a) * growth_factor(cosmo, a) / a


def Gf(cosmo, a):
    """growth.py:124-154: D1' a^3 E."""
    return gp(cosmo, a) * np.power(a, 3) * E(cosmo, a)


def Gf2(cosmo, a):
    """growth.py:157-187: D2' a^3 E."""
    d2f = growth_rate_second(cosmo, a) * growth_factor_second(cosmo, a) / a
    return d2f * np.power(a, 3) * E(cosmo, a)


def dGfa(cosmo, a):
    """growth.py:190-224 (D1'' interpolated in log a as at 221)."""
    t = growth_tables(cosmo)
    d1f = growth_rate(cosmo, a) * growth_factor(cosmo, a) / a
    f1p = np.interp(np.log(a), np.log(t[0]), t[3] / t[0] * t[1])
    ea = E(cosmo, a)
    return f1p * a ** 3 * ea + d1f * a ** 3 * dEa(cosmo, a) + 3 * a ** 2 * ea * d1f


def dGf2a(cosmo, a):
    """growth.py:227-261 (D2'' interpolated in log a as at 258)."""
    t = growth_tables(cosmo)
    d2f = growth_rate_second(cosmo, a) * growth_factor_second(cosmo, a) / a
    f2p = np.interp(np.log(a), np.log(t[0]), t[6] / t[0] * t[4])
    ea = E(cosmo, a)
    return f2p * a ** 3 * ea + d2f * a ** 3 * dEa(cosmo, a) + 3 * a ** 2 * ea * d2f


def Dplusdada(cosmo, a):
    """growth.py:296-314."""
    t = growth_tables(cosmo)
    return np.interp(np.log(np.atleast_1d(a)), np.log(t[0]), t[3] * t[1] / t[0])


def Dplus_to_a(cosmo, D):
    """growth.py:317-334."""
    t = growth_tables(cosmo)
    return np.interp(np.atleast_1d(D), t[1], t[0])
