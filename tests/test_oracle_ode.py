"""Pins for the oracle's ODE siblings (jaxpm/ode.py:13-176): the FastPM
drift and kick factors in Einstein-de Sitter closed form, the exact linear
growth of the FastPM scheme under a linear force law in LCDM (its design
property, for any step size), and the component relations between
make_ode_fn, make_diffrax_ode and symplectic_ode.  A planted mistake in each
factor (on the CPU) must fail the growth pin."""
import numpy as np
import pytest

from oracle import cosmology as oc
from oracle import ode as oo


def eds():
    return oc.Cosmology(Omega_c=0.95, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)


@pytest.mark.parametrize("a,dt0", [(0.1, 0.05), (0.3, 0.1), (0.55, 0.15)])
def test_fpm_factors_eds_closed_form(a, dt0):
    """EdS: D = a, gp = 1, E = a^-1.5, Gf = a^1.5, dGf/da = 1.5 a^0.5, so
    drift = (t0 t1)^-3/4, kick = ((t1 t2)^3/4 - (t0 t1)^3/4) / (1.5 t1 dt0),
    first kick = ((t0 t1)^3/4 - t0^3/2) / (1.5 t0 dt0)."""
    c = eds()
    t0, t1, t2 = a, a + dt0, a + 2 * dt0
    np.testing.assert_allclose(oo.fpm_drift_factor(c, a, dt0), (t0 * t1) ** -0.75, rtol=1e-4)
    np.testing.assert_allclose(oo.fpm_kick_factor(c, a, dt0),
                               ((t1 * t2) ** 0.75 - (t0 * t1) ** 0.75) / (1.5 * t1 * dt0), rtol=1e-4)
    np.testing.assert_allclose(oo.fpm_first_kick_factor(c, a, dt0),
                               ((t0 * t1) ** 0.75 - t0 ** 1.5) / (1.5 * t0 * dt0), rtol=1e-4)


def _linear_growth_error(cosmo, a0, nsteps, dt0, drift=None, kick=None, first_kick=None):
    """Particles displaced by D(a) psi from their lattice sites q, with
    momenta a^3 E D'(a) psi = Gf(a) psi (jaxpm/pm.py:87), under the linear
    force law F(x) = x - q: the growth equation dGf/da = 1.5 Om D / (a^2 E)
    makes x = q + D(a) psi the exact solution, and the FastPM factors
    reproduce it at every step.  Returns the relative position and velocity
    errors at the end."""
    r = np.random.default_rng(0)
    q = r.uniform(0, 32, size=(64, 3))
    psi = r.normal(size=(64, 3))
    x = q + oc.growth_factor(cosmo, a0) * psi
    v = oc.Gf(cosmo, a0) * psi
    forces = lambda pos: pos - q  # noqa: E731
    d, k, fk = oo.symplectic_fpm_ode(None, dt0, cosmo, forces=forces)
    d, k, fk = drift or d, kick or k, first_kick or fk
    v = v + dt0 * fk(a0, x, cosmo)
    a = a0
    ex = ev = 0.0
    for i in range(nsteps):
        x = x + dt0 * d(a, v, cosmo)
        a = a0 + (i + 1) * dt0
        want = oc.growth_factor(cosmo, a) * psi
        ex = max(ex, np.abs(x - q - want).max() / np.abs(want).max())
        if i + 1 < nsteps:  # the growth tables end at a = 1: the last half step is past them
            v = v + dt0 * k(a - dt0, x, cosmo)
            want = oc.Gf(cosmo, np.sqrt(a * (a + dt0))) * psi
            ev = max(ev, np.abs(v - want).max() / np.abs(want).max())
    return ex, ev


@pytest.mark.parametrize("nsteps", [2, 5, 18])
@pytest.mark.parametrize("cosmo", [oc.Cosmology(), oc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7),
                                   oc.Cosmology(Omega_c=0.4, Omega_b=0.05, h=0.6, Omega_k=0.05)])
def test_fastpm_linear_growth_exact(nsteps, cosmo):
    """Exact up to the growth tables' interpolation.  (With w_a != 0 the
    reference's dE/da uses a finite-difference df_de, growth.py:55-85, so
    dGf/da is not the exact derivative and the property does not hold.)"""
    a0 = 0.1
    dt0 = (1.0 - a0) / nsteps
    ex, ev = _linear_growth_error(cosmo, a0, nsteps, dt0)
    assert ex < 5e-5 and ev < 2e-4, (ex, ev)
    # the same driver as a function
    r = np.random.default_rng(0)
    q = r.uniform(0, 32, size=(64, 3))
    psi = r.normal(size=(64, 3))
    x, _ = oo.fastpm_leapfrog(q + oc.growth_factor(cosmo, a0) * psi, oc.Gf(cosmo, a0) * psi, a0, dt0, nsteps,
                              cosmo, None, forces=lambda pos: pos - q)
    np.testing.assert_allclose(x - q, oc.growth_factor(cosmo, 1.0) * psi, rtol=0, atol=2e-5 * np.abs(psi).max())


def test_fastpm_growth_pin_catches_planted_mistakes():
    """Each factor with a plausible mistake (a dropped kick half, the drift
    evaluated at t0 instead of sqrt(t0 t1), the first kick from a to a + dt0
    instead of to the half step) breaks the exact growth."""
    c = oc.Cosmology()
    a0, n = 0.1, 5
    dt0 = (1.0 - a0) / n

    def kick_one_half(a, pos, args):
        t1 = a + dt0
        t0t1 = (a * t1) ** 0.5
        k1 = (oc.Gf(c, t1) - oc.Gf(c, t0t1)) / oc.dGfa(c, t1)
        return 1.0 / (t1 ** 2 * oc.E(c, t1)) * (k1 / dt0) * 1.5 * c.Omega_m * (pos - q)

    def drift_at_t0(a, vel, args):
        return (oc.growth_factor(c, a + dt0) - oc.growth_factor(c, a)) / oc.gp(c, a) / (a ** 3 * oc.E(c, a)) / dt0 * vel

    def first_kick_full(a, pos, args):
        k = (oc.Gf(c, a + dt0) - oc.Gf(c, a)) / oc.dGfa(c, a)
        return 1.0 / (a ** 2 * oc.E(c, a)) * (k / dt0) * 1.5 * c.Omega_m * (pos - q)

    r = np.random.default_rng(0)
    q = r.uniform(0, 32, size=(64, 3))
    for bad in (dict(kick=kick_one_half), dict(drift=drift_at_t0), dict(first_kick=first_kick_full)):
        ex, ev = _linear_growth_error(c, a0, n, dt0, **bad)
        assert max(ex, ev) > 1e-3, (bad, ex, ev)


def test_symplectic_and_diffrax_components_equal_make_ode_fn():
    """ode.py:85-119 and 150-176 are make_ode_fn's two components, split or
    stacked."""
    shape = (8, 8, 8)
    r = np.random.default_rng(1)
    grid = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    x = grid + r.normal(scale=0.5, size=grid.shape)
    v = r.normal(size=grid.shape)
    c = oc.Cosmology()
    a = 0.37
    dpos, dvel = oo.make_ode_fn(shape)((x, v), a, c)
    drift, kick = oo.symplectic_ode(shape, c)
    np.testing.assert_allclose(drift(a, v, c), dpos, rtol=1e-14)
    np.testing.assert_allclose(kick(a, x, c), dvel, rtol=1e-14)
    st = oo.make_diffrax_ode(shape)(a, np.stack([x, v]), c)
    np.testing.assert_allclose(st[0], dpos, rtol=1e-14)
    np.testing.assert_allclose(st[1], dvel, rtol=1e-14)
