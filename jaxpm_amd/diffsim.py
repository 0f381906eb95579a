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

__all__ = ["CosmoCoefficients", "LptNBody", "BatchLptNBody", "SimSpec", "central_jacobian"]


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


class SimSpec:
    """The host side of one simulation of the forward model: its mesh, box,
    time steps, cosmology and the float64 quantities that carry the
    cosmological parameters into the device computation."""

    def __init__(self, mesh_shape, box_size, a_steps, base_cosmo=None, params=("Omega_c", "sigma8"), order=2,
                 k_table=None, rel_step=1e-4):
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
        self.ktab = np.logspace(-4, 1, 128) if k_table is None else np.asarray(k_table, np.float64)
        self.rel_step = rel_step
        S = len(self.a_steps) - 1
        nl = 3 * order
        nt = len(self.ktab)
        self.S = S
        self.slices = {"pk": slice(0, nt), "lpt": slice(nt, nt + nl), "kdk": slice(nt + nl, nt + nl + 1 + 2 * S)}
        self.n = int(np.prod(self.shape))
        self.q = _Quantities(self.quantities, rel_step)

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

    def initial_state(self, noise, params, absolute, lattice, ktab_dev):
        """(q, x0, v0): the host quantities (differentiable in ``params``) and
        the LPT positions and momenta (differentiable in the noise and q)."""
        params = torch.as_tensor(params, dtype=torch.float64) if not isinstance(params, torch.Tensor) else params
        q = CosmoCoefficients.apply(params, self.q, self.rel_step)
        pktab = q[self.slices["pk"]].to(device=noise.device, dtype=torch.float32)
        ic = _autograd.LinearField.apply(noise, self.box, ktab_dev, pktab)
        psi1, psi2 = pm.lpt_fields(ic, lattice if absolute else None, self.order)
        dx, p, _ = pm.lpt_combine(q[self.slices["lpt"]], psi1, psi2)
        x0 = (lattice + dx) if absolute else dx
        return q, x0.reshape(-1, 3), p.reshape(-1, 3)


def _lattice(shape, dev):
    return torch.stack(torch.meshgrid(*[torch.arange(s, device=dev, dtype=torch.float32) for s in shape],
                                      indexing="ij"), -1)


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
    the parameter vector.  ``method="plain"`` runs the same model with the
    leapfrog differentiated by autograd over the stored trajectory
    (:func:`~jaxpm_amd.integrate.leapfrog_autodiff`, no parameter cotangent
    for the leapfrog coefficients), for comparison.
    """

    def __init__(self, mesh_shape, box_size, a_steps, base_cosmo=None, params=("Omega_c", "sigma8"), order=2,
                 paint_absolute_pos=True, checkpoint_every=None, k_table=None, deterministic=False, rel_step=1e-4,
                 method="custom"):
        self.spec = SimSpec(mesh_shape, box_size, a_steps, base_cosmo, params, order, k_table, rel_step)
        self.shape, self.box, self.a_steps = self.spec.shape, self.spec.box, self.spec.a_steps
        self.absolute = bool(paint_absolute_pos)
        self.checkpoint_every = checkpoint_every
        self.deterministic = deterministic
        if method not in ("custom", "plain"):
            raise ValueError(f"unknown method {method!r}")
        self.method = method
        self.stepper = integrate.Stepper(self.shape, self.spec.n, relative=not self.absolute,
                                         deterministic=deterministic) if method == "custom" else None
        self._ktab_dev = None
        self._lattice = None

    def quantities(self, values):
        return self.spec.quantities(values)

    def parameter_vector(self, cosmo=None):
        return self.spec.parameter_vector(cosmo)

    def __call__(self, noise, params):
        noise = to_device(noise)
        if tuple(noise.shape) != self.shape:
            raise ValueError(f"noise shape {tuple(noise.shape)} does not match {self.shape}")
        dev = noise.device
        if self._ktab_dev is None:
            self._ktab_dev = torch.as_tensor(self.spec.ktab, dtype=torch.float32, device=dev)
            self._lattice = _lattice(self.shape, dev)
        q, x0, v0 = self.spec.initial_state(noise, params, self.absolute, self._lattice, self._ktab_dev)
        kdk = q[self.spec.slices["kdk"]]
        if self.method == "custom":
            xS, _ = integrate.NBody.apply(x0, v0, kdk, self.stepper, self.checkpoint_every)
        else:
            xS, _ = integrate.leapfrog_autodiff(x0, v0, kdk, self.shape, self.absolute)
        zeros = torch.zeros(self.shape, dtype=torch.float32, device=dev)
        return _autograd.CicPaint.apply(zeros, xS, None, 1.0, not self.absolute, self.deterministic)


class _BatchLinearField(torch.autograd.Function):
    """Concatenated noise -> concatenated Gaussian fields (symmetric map)."""

    @staticmethod
    def forward(ctx, noise, st, boxes, ktab, pktabs):
        ctx.args = (st, boxes, ktab, pktabs)
        return st.linear_field(noise.contiguous(), boxes, ktab, pktabs)

    @staticmethod
    def backward(ctx, g):
        st, boxes, ktab, pktabs = ctx.args
        return st.linear_field(g.contiguous(), boxes, ktab, pktabs), None, None, None, None


class _BatchLatticeForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, st):
        ctx.st = st
        return st.lattice_forces(field.contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.st.lattice_forces_vjp(g.contiguous()), None


class _BatchScale(torch.autograd.Function):
    """y_b = c[b] x_b with a float32 device array c (jpm_batch_axpy_f32)."""

    @staticmethod
    def forward(ctx, x, st, c):
        ctx.st, ctx.c = st, c
        y = torch.zeros_like(x)
        st.axpy(c, x.contiguous(), y)
        return y

    @staticmethod
    def backward(ctx, g):
        y = torch.zeros_like(g)
        ctx.st.axpy(ctx.c, g.contiguous(), y)
        return y, None, None


class _BatchPaint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, st):
        ctx.st = st
        ctx.save_for_backward(pos)
        return st.paint(pos.contiguous())

    @staticmethod
    def backward(ctx, g):
        (pos,) = ctx.saved_tensors
        return ctx.st.paint_vjp(pos, g.contiguous()), None


class BatchLptNBody:
    """The forward model of :class:`LptNBody` for a batch of simulations with
    their own meshes, boxes and cosmologies, the N-body stage stepped by the
    batched kernels (one launch per stage over all simulations) and, with
    ``graph=True``, replayed from captured HIP graphs.  Simulations of one
    mesh shape must be contiguous.  Calling it with a list of noise meshes
    and a [B, P] parameter tensor returns the list of final densities."""

    def __init__(self, mesh_shapes, box_sizes, a_steps, cosmologies, params=("Omega_c", "sigma8"), order=1,
                 paint_absolute_pos=True, checkpoint_every=None, k_table=None, graph=True, coeff_grads=False,
                 rel_step=1e-4, sort_tuning=None):
        self.specs = [SimSpec(sh, bx, a_steps, c, params, order, k_table, rel_step)
                      for sh, bx, c in zip(mesh_shapes, box_sizes, cosmologies)]
        if not (len(self.specs) == len(mesh_shapes) == len(box_sizes) == len(cosmologies)):
            raise ValueError("one box and one cosmology per mesh shape")
        self.B = len(self.specs)
        self.absolute = bool(paint_absolute_pos)
        self.stepper = integrate.BatchStepper([s.shape for s in self.specs], [s.n for s in self.specs],
                                              relative=not self.absolute, sort_tuning=sort_tuning)
        c0 = np.stack([s.q(s.parameter_vector().numpy())[s.slices["kdk"]] for s in self.specs])
        self.engine = integrate.BatchLeapfrog(self.stepper, c0, checkpoint_every, coeff_grads=coeff_grads,
                                              graph=graph)
        self.checkpoint_every = checkpoint_every
        self._dev_tabs = None

    def parameter_matrix(self):
        return torch.stack([s.parameter_vector() for s in self.specs])

    def _batched_ic(self, dev):
        """Device tables of the batched first-order initial conditions at the
        base cosmologies: boxes, P(k) tables and the two LPT factors, plus the
        concatenated particle lattices."""
        if getattr(self, "_ic", None) is None:
            q = [s.q(s.parameter_vector().numpy()) for s in self.specs]
            boxes = torch.tensor([s.box for s in self.specs], dtype=torch.float32, device=dev)
            ktab = torch.as_tensor(self.specs[0].ktab, dtype=torch.float32, device=dev)
            pk = torch.as_tensor(np.stack([qb[s.slices["pk"]] for qb, s in zip(q, self.specs)]).astype(np.float32),
                                 device=dev)
            lc = np.stack([qb[s.slices["lpt"]] for qb, s in zip(q, self.specs)])
            d1 = torch.as_tensor(lc[:, 0].astype(np.float32), device=dev)
            c1 = torch.as_tensor(lc[:, 1].astype(np.float32), device=dev)
            lat = torch.cat([_lattice(s.shape, dev).reshape(-1, 3) for s in self.specs])
            kdk = torch.as_tensor(np.stack([qb[s.slices["kdk"]] for qb, s in zip(q, self.specs)]))
            self._ic = (boxes, ktab, pk, d1, c1, lat, kdk)
        return self._ic

    def forward_fixed_cosmology(self, noise):
        """The model at the base cosmologies, every stage batched: concatenated
        noise [M] -> concatenated final densities [M].  Differentiable in the
        noise; first-order LPT with absolute painting, one particle per cell
        (what the 48-simulation benchmark runs)."""
        if self.specs[0].order != 1 or not self.absolute:
            raise ValueError("the fully batched model covers 1LPT with absolute painting")
        if any(s.n != int(np.prod(s.shape)) or np.any(s.ktab != self.specs[0].ktab) for s in self.specs):
            raise ValueError("the fully batched model needs one particle per cell and one k table")
        st = self.stepper
        boxes, ktab, pk, d1, c1, lat, kdk = self._batched_ic(noise.device)
        ic = _BatchLinearField.apply(noise, st, boxes, ktab, pk)
        psi = _BatchLatticeForces.apply(ic, st)
        dx = _BatchScale.apply(psi, st, d1)
        p = _BatchScale.apply(dx, st, c1)
        x0 = lat + dx
        xS, _ = integrate.BatchNBody.apply(x0, p, kdk, self.engine, self.checkpoint_every)
        return _BatchPaint.apply(xS, st)

    def __call__(self, noises, params):
        if len(noises) != self.B:
            raise ValueError(f"{len(noises)} noise meshes for {self.B} simulations")
        noises = [to_device(nz) for nz in noises]
        dev = noises[0].device
        if self._dev_tabs is None:
            self._dev_tabs = [(torch.as_tensor(s.ktab, dtype=torch.float32, device=dev), _lattice(s.shape, dev))
                              for s in self.specs]
        params = torch.as_tensor(params, dtype=torch.float64) if not isinstance(params, torch.Tensor) else params
        xs, vs, kdk = [], [], []
        for b, s in enumerate(self.specs):
            kt, lat = self._dev_tabs[b]
            q, x0, v0 = s.initial_state(noises[b], params[b], self.absolute, lat, kt)
            xs.append(x0)
            vs.append(v0)
            kdk.append(q[s.slices["kdk"]])
        x0 = torch.cat(xs)
        v0 = torch.cat(vs)
        xS, _ = integrate.BatchNBody.apply(x0, v0, torch.stack(kdk), self.engine, self.checkpoint_every)
        out, off = [], 0
        for s in self.specs:
            zeros = torch.zeros(s.shape, dtype=torch.float32, device=dev)
            out.append(_autograd.CicPaint.apply(zeros, xS[off:off + s.n], None, 1.0, not self.absolute, False))
            off += s.n
        return out
