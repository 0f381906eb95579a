"""Oracle adjoint of the PM force and the leapfrog (test infrastructure only).

The reference differentiates with ``jax.grad`` (XLA transposes scatter <->
gather and FFT <-> IFFT; tests/test_gradients.py:30-88).  This module writes
the same vector-Jacobian products out in NumPy so that the GPU adjoint can be
compared term by term.  Derivative convention of ``jnp.abs``: d|u|/du = +1 at
u >= 0, -1 below (so the CIC weight 1 - |x - c| has slope -1 on its low
corner and +1 on its high corner).

Pinned by float64 central differences of :mod:`oracle.pm` / :mod:`oracle.ode`
(tests/test_oracle_adjoint.py).
"""
import numpy as np

from . import cosmology as cc
from .kernels import fftk, gradient_kernel, invlaplace_kernel, longrange_kernel
from .painting import ABS_CORNERS, cic_paint
from .pm import force_potentials


def cic_read_grad(mesh, positions):
    """d/dx of cic_read(mesh, x): (n, 3)."""
    mesh = np.asarray(mesh, np.float64)
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    shape = mesh.shape
    fl = np.floor(pos)
    out = np.zeros_like(pos)
    for o in ABS_CORNERS:
        c = fl + o
        u = pos - c
        k = 1.0 - np.abs(u)
        dk = np.where(u >= 0, -1.0, 1.0)
        ci = np.mod(c.astype(np.int64), np.array(shape))
        v = mesh[ci[:, 0], ci[:, 1], ci[:, 2]]
        out[:, 0] += v * dk[:, 0] * k[:, 1] * k[:, 2]
        out[:, 1] += v * k[:, 0] * dk[:, 1] * k[:, 2]
        out[:, 2] += v * k[:, 0] * k[:, 1] * dk[:, 2]
    return out


def force_vjp(positions, cot, mesh_shape, r_split=0.0):
    """J^T cot for F = pm_forces(positions) (absolute painting)."""
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    u = np.asarray(cot, np.float64).reshape(-1, 3)
    rho = cic_paint(np.zeros(mesh_shape), pos)
    phis = force_potentials(np.fft.fftn(rho), r_split)
    g = np.zeros_like(pos)
    for d in range(3):
        g += u[:, d:d + 1] * cic_read_grad(phis[d], pos)
    kvec = fftk(mesh_shape)
    green = invlaplace_kernel(kvec) * longrange_kernel(kvec, r_split)
    rho_bar_k = 0
    for d in range(3):
        gd = cic_paint(np.zeros(mesh_shape), pos, u[:, d])
        # transpose of Re IFFT[H FFT .] is Re IFFT[conj(H) FFT .]
        rho_bar_k = rho_bar_k + np.conj(-gradient_kernel(kvec, d) * green) * np.fft.fftn(gd)
    rho_bar = np.fft.ifftn(rho_bar_k).real
    return g + cic_read_grad(rho_bar, pos)


def kdk_schedule(cosmo, a_steps):
    a = np.asarray(a_steps, np.float64)
    c = 1.5 * cosmo.Omega_m / (a ** 2 * np.sqrt(cc.Esqr(cosmo, a)))
    h = np.diff(a)
    am = a[:-1] + 0.5 * h
    beta = h / (am ** 3 * np.sqrt(cc.Esqr(cosmo, am)))
    gamma = 0.5 * h * c[1:]
    gamma[:-1] += 0.5 * h[1:] * c[1:-1]
    return 0.5 * h[0] * c[0], beta, gamma


def leapfrog_vjp(x0, v0, a_steps, cosmo, mesh_shape, gx, gv):
    """(dL/dx0, dL/dv0) of the KDK leapfrog given (dL/dxS, dL/dvS), storing
    the whole trajectory (no checkpointing)."""
    from .pm import pm_forces
    k0, beta, gamma = kdk_schedule(cosmo, a_steps)
    x = np.asarray(x0, np.float64).reshape(-1, 3).copy()
    v = np.asarray(v0, np.float64).reshape(-1, 3).copy()
    xs = [x.copy()]
    v = v + k0 * pm_forces(x, mesh_shape)
    for i in range(len(beta)):
        x = x + beta[i] * v
        xs.append(x.copy())
        v = v + gamma[i] * pm_forces(x, mesh_shape)
    lx = np.asarray(gx, np.float64).reshape(-1, 3).copy()
    lv = np.asarray(gv, np.float64).reshape(-1, 3).copy()
    for i in reversed(range(len(beta))):
        lx = lx + gamma[i] * force_vjp(xs[i + 1], lv, mesh_shape)
        lv = lv + beta[i] * lx
    lx = lx + k0 * force_vjp(xs[0], lv, mesh_shape)
    return lx, lv
