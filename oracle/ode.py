"""Oracle time derivatives and fixed-step integrators (test infra only).

``make_ode_fn``, ``make_diffrax_ode``, ``symplectic_ode`` and
``symplectic_fpm_ode`` restate JaxPM ``jaxpm/ode.py:13-176``; the FastPM
factors are exposed on their own (``fpm_*_factor``) and pinned in
Einstein-de Sitter closed form and by the exact linear growth of the FastPM
scheme under a linear force law (tests/test_oracle_ode.py).  The reference has
no in-tree fixed-step driver; the build fixes one (kick-drift-kick in the
scale factor ``a`` over the ``make_ode_fn`` right-hand side):

    v += h/2 * dv/da(a, x)
    x += h * v / (am^3 E(am)),  am = a + h/2
    v += h/2 * dv/da(a + h, x)

with ``dv/da = 1.5 Om F(x) / (a^2 E(a))`` (ode.py:137, 143).  This oracle
evaluates one force per half kick (no reuse), so it checks the fused driver's
force reuse as well.
"""
import numpy as np

from . import cosmology as cc
from .painting import cic_paint, cic_paint_dx
from .pm import pm_forces


def make_ode_fn(mesh_shape, paint_absolute_pos=True):
    """ode.py:122-147."""
    def nbody_ode(state, a, cosmo):
        pos, vel = state
        forces = pm_forces(pos, mesh_shape=mesh_shape,
                           paint_absolute_pos=paint_absolute_pos) * 1.5 * cosmo.Omega_m
        e = np.sqrt(cc.Esqr(cosmo, a))
        dpos = 1.0 / (a ** 3 * e) * vel
        dvel = 1.0 / (a ** 2 * e) * forces
        return dpos, dvel
    return nbody_ode


def _pm_accel(mesh_shape, paint_absolute_pos, cosmo, forces):
    """1.5 Omega_m F(pos) (ode.py:46-52); ``forces`` replaces pm_forces when
    given (a callable of the positions), so the integrators can be checked
    against exactly solvable force laws."""
    if forces is None:
        def forces(pos):
            return pm_forces(pos, mesh_shape=mesh_shape, paint_absolute_pos=paint_absolute_pos)
    return lambda pos: forces(pos) * 1.5 * cosmo.Omega_m


def make_diffrax_ode(mesh_shape, paint_absolute_pos=True):
    """ode.py:150-176: ``nbody_ode(a, state, args) -> stack([dpos, dvel])``."""
    inner = make_ode_fn(mesh_shape, paint_absolute_pos)

    def nbody_ode(a, state, args):
        dpos, dvel = inner((state[0], state[1]), a, args)
        return np.stack([dpos, dvel])
    return nbody_ode


def symplectic_ode(mesh_shape, cosmo, paint_absolute_pos=True, forces=None):
    """ode.py:85-119: ``(drift(a, vel, args), kick(a, pos, args))``."""
    accel = _pm_accel(mesh_shape, paint_absolute_pos, cosmo, forces)

    def drift(a, vel, args):
        return 1 / (a ** 3 * cc.E(cosmo, a)) * vel

    def kick(a, pos, args):
        return 1.0 / (a ** 2 * cc.E(cosmo, a)) * accel(pos)
    return drift, kick


def fpm_drift_factor(cosmo, a, dt0):
    """ode.py:20-33: the drift over [a, a + dt0] is factor * vel * dt0, with
    factor = (D(t1) - D(t0)) / gp(ac) / (ac^3 E(ac)) / dt0, ac = sqrt(t0 t1)."""
    t0, t1 = a, a + dt0
    ac = (t0 * t1) ** 0.5
    drift_contr = (cc.growth_factor(cosmo, t1) - cc.growth_factor(cosmo, t0)) / cc.gp(cosmo, ac)
    return 1 / (ac ** 3 * cc.E(cosmo, ac)) * (drift_contr / dt0)


def fpm_kick_factor(cosmo, a, dt0):
    """ode.py:35-58: the kick at t1 = a + dt0 from sqrt(t0 t1) to sqrt(t1 t2)
    is factor * 1.5 Omega_m F * dt0."""
    t0 = a
    t1 = t0 + dt0
    t2 = t1 + dt0
    t0t1 = (t0 * t1) ** 0.5
    t1t2 = (t1 * t2) ** 0.5
    k1 = (cc.Gf(cosmo, t1) - cc.Gf(cosmo, t0t1)) / cc.dGfa(cosmo, t1)
    k2 = (cc.Gf(cosmo, t1t2) - cc.Gf(cosmo, t1)) / cc.dGfa(cosmo, t1)
    return 1.0 / (t1 ** 2 * cc.E(cosmo, t1)) * ((k1 + k2) / dt0)


def fpm_first_kick_factor(cosmo, a, dt0):
    """ode.py:60-80: the first kick from t0 = a to sqrt(t0 t1)."""
    t0 = a
    t1 = t0 + dt0
    t0t1 = (t0 * t1) ** 0.5
    k = (cc.Gf(cosmo, t0t1) - cc.Gf(cosmo, t0)) / cc.dGfa(cosmo, t0)
    return 1.0 / (a ** 2 * cc.E(cosmo, a)) * (k / dt0)


def symplectic_fpm_ode(mesh_shape, dt0, cosmo, paint_absolute_pos=True, forces=None):
    """ode.py:13-82: FastPM ``(drift, kick, first_kick)``; each returns the
    rate whose product with dt0 is the step's update."""
    accel = _pm_accel(mesh_shape, paint_absolute_pos, cosmo, forces)

    def drift(a, vel, args):
        return fpm_drift_factor(cosmo, a, dt0) * vel

    def kick(a, pos, args):
        return fpm_kick_factor(cosmo, a, dt0) * accel(pos)

    def first_kick(a, pos, args):
        return fpm_first_kick_factor(cosmo, a, dt0) * accel(pos)
    return drift, kick, first_kick


def fastpm_leapfrog(pos, vel, a0, dt0, nsteps, cosmo, mesh_shape, paint_absolute_pos=True, forces=None):
    """The FastPM scheme driven by ``symplectic_fpm_ode`` as a symplectic
    solver steps it: one first kick at a0 (velocities to sqrt(a0 a1)), then
    per step a drift over [a, a + dt0] and a kick at a + dt0 (velocities to
    the next half step).  Returns positions at a0 + nsteps dt0 and velocities
    at sqrt(a_n a_{n+1})."""
    drift, kick, first_kick = symplectic_fpm_ode(mesh_shape, dt0, cosmo, paint_absolute_pos, forces)
    x = np.asarray(pos, np.float64)
    v = np.asarray(vel, np.float64)
    v = v + dt0 * first_kick(a0, x, cosmo)
    a = a0
    for _ in range(nsteps):
        x = x + dt0 * drift(a, v, cosmo)
        v = v + dt0 * kick(a, x, cosmo)
        a = a + dt0
    return x, v


def leapfrog(pos, vel, a_steps, cosmo, mesh_shape, paint_absolute_pos=True, reuse_force=False):
    """Fixed-step KDK over the scale factors ``a_steps`` (a0 < a1 < ...).
    By default each half kick evaluates its own force; ``reuse_force`` takes
    the force at a step's end for the next step's first half kick (the same
    positions, so the same values), which halves the work."""
    x = np.asarray(pos, np.float64)
    v = np.asarray(vel, np.float64)

    def accel(x, a):
        f = pm_forces(x, mesh_shape=mesh_shape, paint_absolute_pos=paint_absolute_pos) * 1.5 * cosmo.Omega_m
        return f, 1.0 / (a ** 2 * np.sqrt(cc.Esqr(cosmo, a)))

    ode = make_ode_fn(mesh_shape, paint_absolute_pos)
    f = None
    for a0, a1 in zip(a_steps[:-1], a_steps[1:]):
        h = a1 - a0
        if reuse_force:
            if f is None:
                f, _ = accel(x, a0)
            dv = f / (a0 ** 2 * np.sqrt(cc.Esqr(cosmo, a0)))
        else:
            _, dv = ode((x, v), a0, cosmo)
        v = v + 0.5 * h * dv
        am = a0 + 0.5 * h
        x = x + h * v / (am ** 3 * np.sqrt(cc.Esqr(cosmo, am)))
        if reuse_force:
            f, _ = accel(x, a1)
            dv = f / (a1 ** 2 * np.sqrt(cc.Esqr(cosmo, a1)))
        else:
            _, dv = ode((x, v), a1, cosmo)
        v = v + 0.5 * h * dv
    return x, v


def final_density(pos, mesh_shape, paint_absolute_pos=True):
    if paint_absolute_pos:
        return cic_paint(np.zeros(mesh_shape), pos)
    return cic_paint_dx(pos)
