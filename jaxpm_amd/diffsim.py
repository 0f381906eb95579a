"""Differentiable simulations: gradients of a loss on the final density with
respect to the white-noise field and to cosmological parameters.

The forward model is the one the reference differentiates in
tests/test_gradients.py:27-72 -- ``linear_field`` -> ``lpt`` -> N-body ->
``cic_paint`` -- with the checkpointed leapfrog of :mod:`jaxpm_amd.integrate`
in place of diffrax.  Every device step has a hand-written adjoint
(``_autograd`` and :class:`~jaxpm_amd.integrate.NBody`).

Cosmology enters the device computation only through host-side float64
quantities: the P(k) table that ``linear_field`` interpolates, the LPT growth
coefficients and the leapfrog coefficients.  :class:`CosmoCoefficients`
evaluates them and, in the backward pass, closes the chain rule on the host:
the device returns the cotangent of each quantity (P(k)-table cotangents,
float64 dot products for the scalars) and the Jacobian of the quantities with
respect to the parameters is taken by central differences of the host code.
That code is smooth in the parameters -- fixed-step RK4 growth tables,
closed-form Eisenstein-Hu, fixed-grid Simpson for sigma8 -- so the
differences carry no solver noise.  (jax.grad differentiates the same
quantities through jax_cosmo's ODE and fitting formulas.)
"""
import numpy as np
import torch

from . import _autograd, integrate, pm
from . import cosmology as cs
from ._convert import to_device

__all__ = ["CosmoCoefficients", "LptNBody", "central_jacobian"]


def central_jacobian(fn, x, rel_step):
    """(len(fn(x)), len(x)) central-difference Jacobian of a host function,
    step ``rel_step * max(|x_j|, 1e-3)`` per parameter."""
    cols = []
    for j in range(len(x)):
        h = rel_step * max(abs(x[j]), 1e-3)
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        cols.append((np.asarray(fn(xp), np.float64) - np.asarray(fn(xm), np.float64)) / (2 * h))
    return np.stack(cols, 1)


class CosmoCoefficients(torch.autograd.Function):
    """params (float64 host vector) -> fn(params) (float64 host vector).

    Backward: J^T g with J = ``fn.jacobian(x)`` when ``fn`` provides one (a
    cached Jacobian), else the central-difference Jacobian of ``fn``."""

    @staticmethod
    def forward(ctx, params, fn, rel_step):
        x = params.detach().to(torch.float64).cpu().numpy().copy()
        ctx.fn, ctx.x, ctx.rel_step = fn, x, float(rel_step)
        ctx.meta = (params.dtype, params.device)
        return torch.as_tensor(np.asarray(fn(x), np.float64), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        jac = getattr(ctx.fn, "jacobian", None)
        J = jac(ctx.x) if jac is not None else central_jacobian(ctx.fn, ctx.x, ctx.rel_step)
        gv = g.detach().to(torch.float64).cpu().numpy()
        return torch.as_tensor(J.T @ gv, dtype=ctx.meta[0], device=ctx.meta[1]), None, None


class _Quantities:
    """A host function of the parameter vector with its value and Jacobian
    memoised per parameter point (a noise-only optimisation reuses them)."""

    def __init__(self, fn, rel_step, size=8):
        self.fn, self.rel_step, self.size = fn, rel_step, size
        self._val, self._jac = {}, {}

    def _put(self, d, k, v):
        if len(d) >= self.size:
            d.pop(next(iter(d)))
        d[k] = v
        return v

    def __call__(self, x):
        k = tuple(float(v) for v in x)
        return self._val[k] if k in self._val else self._put(self._val, k, self.fn(np.asarray(x, np.float64)))

    def jacobian(self, x):
        k = tuple(float(v) for v in x)
        if k in self._jac:
            return self._jac[k]
        return self._put(self._jac, k, central_jacobian(self.fn, np.asarray(x, np.float64), self.rel_step))


class LptNBody:
    """noise [nx,ny,nz], cosmological parameters -> final CIC density.

    ``params`` names the fields of :class:`~jaxpm_amd.cosmology.Cosmology`
    that vary (default Omega_c and sigma8); the others come from
    ``base_cosmo``.  The model is linear_field with the Eisenstein-Hu table on
    ``k_table`` (jnp.interp, as conftest.py:90-94), ``order``-LPT at ``a0``
    onto the particle lattice, the leapfrog over ``a_steps`` (starting at
    ``a0``) with the adjoint checkpointed every ``checkpoint_every`` steps,
    and the CIC paint of the final positions.  Calling it returns the density
    mesh; torch autograd gives its cotangent with respect to the noise and
    the parameter vector.
    """

    def __init__(self, mesh_shape, box_size, a_steps, base_cosmo=None, params=("Omega_c", "sigma8"), order=2,
                 paint_absolute_pos=True, checkpoint_every=None, k_table=None, deterministic=False, rel_step=1e-4):
        self.shape = tuple(int(s) for s in mesh_shape)
        self.box = tuple(float(b) for b in box_size)
        self.a_steps = np.asarray(a_steps, np.float64)
        self.a0 = float(self.a_steps[0])
        self.base = base_cosmo if base_cosmo is not None else cs.Cosmology(
            Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
        self.names = tuple(params)
        for nm in self.names:
            if not hasattr(self.base, nm):
                raise ValueError(f"unknown cosmological parameter {nm!r}")
        if order not in (1, 2):
            raise ValueError(f"order must be 1 or 2, got {order}")
        self.order = order
        self.absolute = bool(paint_absolute_pos)
        self.checkpoint_every = checkpoint_every
        self.ktab = np.logspace(-4, 1, 128) if k_table is None else np.asarray(k_table, np.float64)
        self.deterministic = deterministic
        self.rel_step = rel_step
        S = len(self.a_steps) - 1
        nl = 3 * order
        nt = len(self.ktab)
        self.slices = {"pk": slice(0, nt), "lpt": slice(nt, nt + nl), "kdk": slice(nt + nl, nt + nl + 1 + 2 * S)}
        n = int(np.prod(self.shape))
        self.stepper = integrate.Stepper(self.shape, n, relative=not self.absolute, deterministic=deterministic)
        self._ktab_dev = None
        self._lattice = None
        self._q = _Quantities(self.quantities, rel_step)

    def cosmology(self, values):
        return self.base.replace(**{nm: float(v) for nm, v in zip(self.names, values)})

    def quantities(self, values):
        """Host float64 vector [P(k) table, LPT coefficients, leapfrog
        coefficients] at the given parameter values."""
        c = self.cosmology(values)
        pk = cs.linear_matter_power(c, self.ktab)
        lc = pm.lpt_coefficients(c, self.a0, self.order)
        k0, beta, gamma = integrate.kdk_schedule(c, self.a_steps)
        return np.concatenate([pk, lc, [k0], beta, gamma])

    def parameter_vector(self, cosmo=None):
        c = self.base if cosmo is None else cosmo
        return torch.tensor([float(getattr(c, nm)) for nm in self.names], dtype=torch.float64)

    def __call__(self, noise, params):
        noise = to_device(noise)
        if tuple(noise.shape) != self.shape:
            raise ValueError(f"noise shape {tuple(noise.shape)} does not match {self.shape}")
        params = torch.as_tensor(params, dtype=torch.float64) if not isinstance(params, torch.Tensor) else params
        q = CosmoCoefficients.apply(params, self._q, self.rel_step)
        dev = noise.device
        if self._ktab_dev is None:
            self._ktab_dev = torch.as_tensor(self.ktab, dtype=torch.float32, device=dev)
        pktab = q[self.slices["pk"]].to(device=dev, dtype=torch.float32)
        ic = _autograd.LinearField.apply(noise, self.box, self._ktab_dev, pktab)
        if self.absolute:
            if self._lattice is None:
                self._lattice = torch.stack(torch.meshgrid(
                    *[torch.arange(s, device=dev, dtype=torch.float32) for s in self.shape], indexing="ij"), -1)
            particles = self._lattice
        else:
            particles = None
        psi1, psi2 = pm.lpt_fields(ic, particles, self.order)
        dx, p, _ = pm.lpt_combine(q[self.slices["lpt"]], psi1, psi2)
        x0 = (particles + dx) if self.absolute else dx
        xS, _ = integrate.NBody.apply(x0.reshape(-1, 3), p.reshape(-1, 3), q[self.slices["kdk"]], self.stepper,
                                      self.checkpoint_every)
        zeros = torch.zeros(self.shape, dtype=torch.float32, device=dev)
        return _autograd.CicPaint.apply(zeros, xS, None, 1.0, not self.absolute, self.deterministic)
