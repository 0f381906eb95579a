"""Oracle PM force, LPT and Gaussian initial conditions (test infra only).

Restates JaxPM ``jaxpm/pm.py:12-144`` in float64 with complex-to-complex
``numpy.fft`` (``fft3d``/``ifft3d`` of ``distributed.py:37-42``: unnormalised
forward, inverse with 1/M followed by ``.real``).
"""
import numpy as np

from . import cosmology as cc
from .kernels import fftk, gradient_kernel, invlaplace_kernel, longrange_kernel
from .painting import abs_corners, cic_paint, cic_paint_dx, cic_read, cic_read_dx


def fft3d(x):
    return np.fft.fftn(np.asarray(x))


def ifft3d(x):
    return np.fft.ifftn(x).real


def uniform_particles(mesh_shape):
    """distributed.py:187-190: integer lattice (nx, ny, nz, 3)."""
    return np.stack(np.meshgrid(*[np.arange(s) for s in mesh_shape], indexing='ij'),
                    axis=-1).astype(np.int32)


def force_potentials(delta_k, r_split=0.0, order=1):
    """The three real meshes Re IFFT[-g_d G Lr delta_k] of pm.py:49-56."""
    shape = delta_k.shape
    kvec = fftk(shape)
    pot_k = delta_k * invlaplace_kernel(kvec) * longrange_kernel(kvec, r_split)
    return [ifft3d(-gradient_kernel(kvec, i, order) * pot_k) for i in range(3)]


def pm_forces(positions, mesh_shape=None, delta=None, r_split=0,
              paint_absolute_pos=True):
    """pm.py:12-58."""
    if mesh_shape is None:
        assert delta is not None
        mesh_shape = delta.shape
    if delta is None:
        if paint_absolute_pos:
            field = cic_paint(np.zeros(mesh_shape), positions)
        else:
            field = cic_paint_dx(positions)
        delta_k = fft3d(field)
    elif np.isrealobj(delta):
        delta_k = fft3d(delta)
    else:
        delta_k = np.asarray(delta)
    pots = force_potentials(delta_k, r_split)
    if paint_absolute_pos:
        corners = abs_corners(positions, pots[0].shape)
        comps = [cic_read(p, positions, corners=corners) for p in pots]
    else:
        comps = [cic_read_dx(p, positions) for p in pots]
    return np.stack(comps, axis=-1)


def lpt(cosmo, initial_conditions, particles=None, a=0.1, order=1):
    """pm.py:61-126; returns (dx, p, f)."""
    ic = np.asarray(initial_conditions, np.float64)
    absolute = particles is not None
    if particles is None:
        particles = np.zeros((*ic.shape, 3))
    a = np.atleast_1d(np.float64(a))
    E = np.sqrt(cc.Esqr(cosmo, a))
    delta_k = fft3d(ic)
    f1 = pm_forces(particles, delta=delta_k, paint_absolute_pos=absolute)
    dx = cc.growth_factor(cosmo, a) * f1
    p = a ** 2 * cc.growth_rate(cosmo, a) * E * dx
    f = a ** 2 * E * cc.dGfa(cosmo, a) * f1
    if order == 2:
        delta2 = second_order_source(delta_k)
        f2 = pm_forces(particles, delta=fft3d(delta2), paint_absolute_pos=absolute)
        dx2 = 3 / 7 * cc.growth_factor_second(cosmo, a) * f2
        p2 = a ** 2 * cc.growth_rate_second(cosmo, a) * E * dx2
        ff2 = a ** 2 * E * cc.dGf2a(cosmo, a) * f2
        dx = dx + dx2
        p = p + p2
        f = f + ff2
    return dx, p, f


def second_order_source(delta_k):
    """pm.py:89-111: sum_{i<j} (phi_ii phi_jj - phi_ij^2) in the reference's
    accumulation order."""
    kvec = fftk(delta_k.shape)
    pot_k = delta_k * invlaplace_kernel(kvec)
    delta2 = 0
    shear_acc = 0
    for i in range(3):
        s_ii = ifft3d(gradient_kernel(kvec, i) ** 2 * pot_k)
        delta2 = delta2 + s_ii * shear_acc
        shear_acc = shear_acc + s_ii
        for j in range(i + 1, 3):
            s_ij = ifft3d(gradient_kernel(kvec, i) * gradient_kernel(kvec, j) * pot_k)
            delta2 = delta2 - s_ij ** 2
    return delta2


def linear_field(mesh_shape, box_size, pk, noise):
    """pm.py:129-144 with the white noise supplied as an array (``noise``
    replaces ``jax.random.normal(seed)`` of distributed.py:222-223)."""
    field = fft3d(np.asarray(noise, np.float64))
    kvec = fftk(mesh_shape)
    kmesh = sum((kk / box_size[i] * mesh_shape[i]) ** 2 for i, kk in enumerate(kvec)) ** 0.5
    pkmesh = pk(kmesh) * (mesh_shape[0] * mesh_shape[1] * mesh_shape[2]) / (
        box_size[0] * box_size[1] * box_size[2])
    return ifft3d(field * np.sqrt(pkmesh))
