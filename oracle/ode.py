"""Oracle time derivative and fixed-step KDK leapfrog (test infra only).

``make_ode_fn`` restates JaxPM ``jaxpm/ode.py:122-147``.  The reference has
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


def leapfrog(pos, vel, a_steps, cosmo, mesh_shape, paint_absolute_pos=True):
    """Fixed-step KDK over the scale factors ``a_steps`` (a0 < a1 < ...)."""
    ode = make_ode_fn(mesh_shape, paint_absolute_pos)
    x = np.asarray(pos, np.float64)
    v = np.asarray(vel, np.float64)
    for a0, a1 in zip(a_steps[:-1], a_steps[1:]):
        h = a1 - a0
        _, dv = ode((x, v), a0, cosmo)
        v = v + 0.5 * h * dv
        am = a0 + 0.5 * h
        x = x + h * v / (am ** 3 * np.sqrt(cc.Esqr(cosmo, am)))
        _, dv = ode((x, v), a1, cosmo)
        v = v + 0.5 * h * dv
    return x, v


def final_density(pos, mesh_shape, paint_absolute_pos=True):
    if paint_absolute_pos:
        return cic_paint(np.zeros(mesh_shape), pos)
    return cic_paint_dx(pos)
