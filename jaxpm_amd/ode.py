"""Time-derivative builders, mirroring ``jaxpm/ode.py``.

Each right-hand side calls :func:`jaxpm_amd.pm.pm_forces` (one fused force
evaluation on the GPU) and host-side scalar coefficients.  The fixed-step
kick-drift-kick driver with a checkpointed adjoint lives in
:mod:`jaxpm_amd.integrate`.
"""
import numpy as np
import torch

from . import growth as gr
from .pm import pm_forces

__all__ = ["make_ode_fn", "make_diffrax_ode", "symplectic_ode", "symplectic_fpm_ode"]


def _scalar(a):
    if isinstance(a, torch.Tensor):
        return float(a.detach().reshape(-1)[0].item())
    return float(np.asarray(a, np.float64).reshape(-1)[0])


def make_ode_fn(mesh_shape, paint_absolute_pos=True, halo_size=0, sharding=None):
    """ode.py:122-147: ``nbody_ode((pos, vel), a, cosmo) -> (dpos, dvel)``."""

    def nbody_ode(state, a, cosmo):
        pos, vel = state
        a = _scalar(a)
        forces = pm_forces(pos, mesh_shape=mesh_shape, paint_absolute_pos=paint_absolute_pos,
                           halo_size=halo_size, sharding=sharding) * (1.5 * cosmo.Omega_m)
        e = float(np.sqrt(gr.Esqr(cosmo, a)))
        dpos = (1.0 / (a ** 3 * e)) * vel
        dvel = (1.0 / (a ** 2 * e)) * forces
        return dpos, dvel

    return nbody_ode


def make_diffrax_ode(mesh_shape, paint_absolute_pos=True, halo_size=0, sharding=None):
    """ode.py:150-176: ``nbody_ode(a, state, args) -> stack([dpos, dvel])``."""
    inner = make_ode_fn(mesh_shape, paint_absolute_pos, halo_size, sharding)

    def nbody_ode(a, state, args):
        dpos, dvel = inner((state[0], state[1]), a, args)
        return torch.stack([dpos, dvel])

    return nbody_ode


def symplectic_ode(mesh_shape, cosmo, paint_absolute_pos=True, halo_size=0, sharding=None):
    """ode.py:85-119: ``(drift(a, vel, args), kick(a, pos, args))``."""

    def drift(a, vel, args):
        a = _scalar(a)
        return (1.0 / (a ** 3 * float(gr.E(cosmo, a)))) * vel

    def kick(a, pos, args):
        a = _scalar(a)
        forces = pm_forces(pos, mesh_shape=mesh_shape, paint_absolute_pos=paint_absolute_pos,
                           halo_size=halo_size, sharding=sharding) * (1.5 * cosmo.Omega_m)
        return (1.0 / (a ** 2 * float(gr.E(cosmo, a)))) * forces

    return drift, kick


def symplectic_fpm_ode(mesh_shape, dt0, cosmo, paint_absolute_pos=True, halo_size=0, sharding=None):
    """ode.py:13-82: FastPM-factor ``(drift, kick, first_kick)``."""

    def drift(a, vel, args):
        t0 = _scalar(a)
        t1 = t0 + dt0
        ac = (t0 * t1) ** 0.5
        drift_contr = (gr.growth_factor(cosmo, t1) - gr.growth_factor(cosmo, t0)) / gr.gp(cosmo, ac)
        f = 1.0 / (ac ** 3 * gr.E(cosmo, ac)) * drift_contr / dt0
        return float(f) * vel

    def _forces(pos):
        return pm_forces(pos, mesh_shape=mesh_shape, paint_absolute_pos=paint_absolute_pos,
                         halo_size=halo_size, sharding=sharding) * (1.5 * cosmo.Omega_m)

    def kick(a, pos, args):
        t0 = _scalar(a)
        t1 = t0 + dt0
        t2 = t1 + dt0
        t0t1 = (t0 * t1) ** 0.5
        t1t2 = (t1 * t2) ** 0.5
        ac = t1
        k1 = (gr.Gf(cosmo, t1) - gr.Gf(cosmo, t0t1)) / gr.dGfa(cosmo, t1)
        k2 = (gr.Gf(cosmo, t1t2) - gr.Gf(cosmo, t1)) / gr.dGfa(cosmo, t1)
        f = 1.0 / (ac ** 2 * gr.E(cosmo, ac)) * (k1 + k2) / dt0
        return float(f) * _forces(pos)

    def first_kick(a, pos, args):
        t0 = _scalar(a)
        t1 = t0 + dt0
        t0t1 = (t0 * t1) ** 0.5
        k = (gr.Gf(cosmo, t0t1) - gr.Gf(cosmo, t0)) / gr.dGfa(cosmo, t0)
        f = 1.0 / (t0 ** 2 * gr.E(cosmo, t0)) * k / dt0
        return float(f) * _forces(pos)

    return drift, kick, first_kick
