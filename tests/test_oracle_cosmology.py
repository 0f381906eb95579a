"""Pins for the oracle cosmology (textbook limits, Heath integral) and the
product's host cosmology against it."""
import numpy as np
import pytest

from oracle import cosmology as oc


def eds():
    return oc.Cosmology(Omega_c=0.95, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)


def test_eds_limits():
    c = eds()
    a = np.array([0.05, 0.1, 0.3, 0.7, 1.0])
    np.testing.assert_allclose(oc.growth_factor(c, a), a, rtol=2e-6)
    np.testing.assert_allclose(oc.growth_rate(c, a), 1.0, rtol=1e-6)
    # D2 = a^2 is curved: exact at the table nodes, and between nodes off by
    # at most the linear-interpolation bound (da)^2 / 4.
    nodes = oc.growth_tables(c)[0][::15]
    np.testing.assert_allclose(oc.growth_factor_second(c, nodes), nodes ** 2, rtol=1e-7)
    da = a * (10 ** (3.0 / 255) - 1)
    assert np.all(np.abs(oc.growth_factor_second(c, a) - a ** 2) <= da ** 2 / 4 * 1.01 + 1e-12)
    np.testing.assert_allclose(oc.growth_rate_second(c, a), 2.0, rtol=1e-5)
    np.testing.assert_allclose(oc.E(c, a), a ** -1.5, rtol=1e-12)
    # Gf = D' a^3 E = a^1.5 in EdS;  dGf/da = 1.5 a^0.5
    np.testing.assert_allclose(oc.Gf(c, a), a ** 1.5, rtol=2e-6)
    np.testing.assert_allclose(oc.dGfa(c, a), 1.5 * a ** 0.5, rtol=1e-5)


@pytest.mark.parametrize("om", [0.25, 0.3111])
def test_lcdm_growth_matches_heath_integral(om):
    c = oc.Cosmology(Omega_c=om - 0.049, Omega_b=0.049)
    atab = oc.growth_tables(c)[0]
    a = atab[::17]
    heath = np.array([oc.heath_growth(c, x) for x in a]) / oc.heath_growth(c, 1.0)
    np.testing.assert_allclose(oc.growth_factor(c, a), heath, rtol=2e-7)


def test_eisenstein_hu_large_scale_limit():
    c = oc.Cosmology()
    assert abs(oc.eisenstein_hu(c, 1e-5) - 1.0) < 1e-3
    t = oc.eisenstein_hu(c, np.logspace(-3, 1, 50))
    assert np.all(np.diff(t) < 1e-3)  # broadly decreasing with wiggles


def test_product_cosmology_matches_oracle():
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import growth as pg
    for kw in [dict(), dict(Omega_c=0.2607, Omega_b=0.049, h=0.6766, n_s=0.9665, sigma8=0.8102),
               dict(Omega_c=0.3, w0=-0.9, wa=0.1)]:
        o = oc.Cosmology(**kw)
        p = pc.Cosmology(**kw)
        ot, pt = oc.growth_tables(o), pc.growth_tables(p)
        for a, b in zip(ot, pt):
            np.testing.assert_allclose(b, a, rtol=1e-7, atol=1e-9)
        a = np.array([0.1, 0.2, 0.55, 1.0])
        np.testing.assert_allclose(pg.E(p, a), oc.E(o, a), rtol=1e-13)
        np.testing.assert_allclose(pg.dGfa(p, a), oc.dGfa(o, a), rtol=1e-8)
        np.testing.assert_allclose(pg.dGf2a(p, a), oc.dGf2a(o, a), rtol=1e-8)
        np.testing.assert_allclose(pg.Gf(p, a), oc.Gf(o, a), rtol=1e-8)
        np.testing.assert_allclose(pg.gp(p, a), oc.gp(o, a), rtol=1e-8)
        k = np.logspace(-4, 1, 64)
        np.testing.assert_allclose(pc.eisenstein_hu(p, k), oc.eisenstein_hu(o, k), rtol=1e-12)
        np.testing.assert_allclose(pc.linear_matter_power(p, k), oc.linear_matter_power(o, k), rtol=1e-6)
