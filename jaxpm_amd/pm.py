"""PM force, LPT and Gaussian initial conditions, mirroring ``jaxpm/pm.py``.

``pm_forces`` runs the single-call pipeline ``jpm_pm_forces_f32`` (paint, R2C,
fused k-space pass, batched C2R, 3-component read) whenever no gradient is
requested, and otherwise composes the same kernels through the hand-written
adjoints of ``_autograd``.
"""
import numpy as np
import torch

from . import _autograd, _ops
from . import growth as gr
from ._convert import forward_only, single_device, to_device, wants_f64
from .distributed import normal_field

__all__ = ["pm_forces", "lpt", "linear_field", "TabulatedPk", "lpt_coefficients", "lpt_fields", "lpt_combine"]


def _needs_grad(*ts):
    return torch.is_grad_enabled() and any(isinstance(t, torch.Tensor) and t.requires_grad for t in ts)


def _forces_from_field(field, positions, relative, r_split, kf, scale=1.0, deterministic=False):
    """read(Re IFFT[-g_d G Lr FFT(field)], positions) * scale, (n, 3)."""
    shape = tuple(field.shape)
    if _needs_grad(field, positions):
        phi = _autograd.ForceField.apply(field, r_split, kf)
        f = _autograd.CicRead.apply(phi, positions.reshape(-1, 3), relative, deterministic)
        return f * scale if scale != 1.0 else f
    return _ops.pm_forces_fused(positions, shape, delta_k=_ops.r2c(field), r_split=r_split, scale=scale,
                                relative=relative, kf=kf)


def pm_forces(positions, mesh_shape=None, delta=None, r_split=0, paint_absolute_pos=True, halo_size=0,
              sharding=None, deterministic=False, grad_order=1, lapl_fd=False, compensate=False):
    """pm.py:12-58: gravitational forces on particles, ``positions.shape``.

    ``delta`` may be a real density mesh or its (c2c) spectrum, in which case
    painting is skipped (pm.py:41-47).  Additive keywords: ``deterministic``
    (reproducible paint), ``grad_order`` (0 = exact ``i k``), ``lapl_fd``
    (``invlaplace_kernel(fd=True)``), ``compensate`` (CIC deconvolution inside
    the k-space pass; off by default, as in the reference).
    """
    single_device(sharding)
    if mesh_shape is None:
        assert delta is not None, "If mesh_shape is not provided, delta should be provided"
        mesh_shape = tuple(delta.shape)
    mesh_shape = tuple(int(s) for s in mesh_shape)
    rel = not paint_absolute_pos
    kf = _ops.kflags(grad_order, lapl_fd, compensate)
    if wants_f64(positions, delta):
        pos = to_device(positions, torch.float64)
        forward_only(pos, delta)
        field = None
        if delta is not None:
            if delta.is_complex() if isinstance(delta, torch.Tensor) else np.iscomplexobj(delta):
                raise NotImplementedError("the float64 pm_forces takes a real density (delta), not its spectrum")
            field = to_device(delta, torch.float64)
        return _ops.pm_forces_f64(pos.reshape(-1, 3), mesh_shape, field, float(r_split), rel,
                                  kf).reshape(pos.shape)
    pos = to_device(positions)
    out_shape = pos.shape
    pflat = pos.reshape(-1, 3)
    if delta is None:
        if _needs_grad(pos):
            if rel:
                zeros = torch.zeros(mesh_shape, dtype=torch.float32, device=pos.device)
                rho = _autograd.CicPaint.apply(zeros, pflat, None, 1.0, True, deterministic)
            else:
                zeros = torch.zeros(mesh_shape, dtype=torch.float32, device=pos.device)
                rho = _autograd.CicPaint.apply(zeros, pflat, None, 1.0, False, deterministic)
            phi = _autograd.ForceField.apply(rho, float(r_split), kf)
            return _autograd.CicRead.apply(phi, pflat, rel, deterministic).reshape(out_shape)
        return _ops.pm_forces_fused(pflat, mesh_shape, r_split=float(r_split), relative=rel,
                                    deterministic=deterministic, kf=kf).reshape(out_shape)
    is_complex = delta.is_complex() if isinstance(delta, torch.Tensor) else np.iscomplexobj(delta)
    if is_complex:
        dk = _ops.full_to_half(to_device(delta, torch.complex64))
        return _ops.pm_forces_fused(pflat, mesh_shape, delta_k=dk, r_split=float(r_split), relative=rel,
                                    kf=kf).reshape(out_shape)
    field = to_device(delta)
    return _forces_from_field(field, pflat, rel, float(r_split), kf, deterministic=deterministic).reshape(out_shape)


def _coef(x):
    return float(np.asarray(x, np.float64).reshape(-1)[0])


def lpt(cosmo, initial_conditions, particles=None, a=0.1, halo_size=0, sharding=None, order=1):
    """pm.py:61-126: first/second order LPT displacement, momentum and force
    ``(dx, p, f)``, each ``(nx, ny, nz, 3)``.  ``particles=None`` selects the
    relative (lattice + displacement) read."""
    single_device(sharding)
    if order not in (1, 2):
        raise ValueError(f"order must be 1 or 2, got {order}")
    ic = to_device(initial_conditions)
    shape = tuple(ic.shape)
    rel = particles is None
    if rel:
        parts = torch.zeros((*shape, 3), dtype=torch.float32, device=ic.device)
    else:
        parts = to_device(particles)
    a = _coef(a)
    E = _coef(np.sqrt(gr.Esqr(cosmo, a)))
    f1 = _forces_from_field(ic, parts.reshape(-1, 3), rel, 0.0, 0).reshape(parts.shape)
    d1 = _coef(gr.growth_factor(cosmo, a))
    dx = d1 * f1
    p = (a ** 2 * _coef(gr.growth_rate(cosmo, a)) * E) * dx
    f = (a ** 2 * E * _coef(gr.dGfa(cosmo, a))) * f1
    if order == 2:
        phi6 = _autograd.ShearField.apply(ic, 0) if _needs_grad(ic) else \
            _ops.c2r(_ops.kspace_shear6(_ops.r2c(ic), shape, 1.0 / ic.numel()), shape, clobber=True)
        delta2 = _autograd.Lpt2Source.apply(phi6) if _needs_grad(phi6) else _ops.lpt2_source(phi6)
        f2 = _forces_from_field(delta2, parts.reshape(-1, 3), rel, 0.0, 0).reshape(parts.shape)
        dx2 = (3 / 7 * _coef(gr.growth_factor_second(cosmo, a))) * f2
        p2 = (a ** 2 * _coef(gr.growth_rate_second(cosmo, a)) * E) * dx2
        ff2 = (a ** 2 * E * _coef(gr.dGf2a(cosmo, a))) * f2
        dx = dx + dx2
        p = p + p2
        f = f + ff2
    return dx, p, f


def lpt_coefficients(cosmo, a=0.1, order=1):
    """Host float64 growth coefficients of :func:`lpt` (pm.py:77-124):
    [D1, a^2 f1 E, a^2 E dGfa] and, for order 2, [3/7 D2, a^2 f2 E,
    a^2 E dGf2a] -- dx1 = c0 Psi1, p1 = c1 dx1, f1 = c2 Psi1 and likewise for
    the second-order field Psi2."""
    a = _coef(a)
    E = _coef(np.sqrt(gr.Esqr(cosmo, a)))
    c = [_coef(gr.growth_factor(cosmo, a)), a ** 2 * _coef(gr.growth_rate(cosmo, a)) * E,
         a ** 2 * E * _coef(gr.dGfa(cosmo, a))]
    if order == 2:
        c += [3 / 7 * _coef(gr.growth_factor_second(cosmo, a)), a ** 2 * _coef(gr.growth_rate_second(cosmo, a)) * E,
              a ** 2 * E * _coef(gr.dGf2a(cosmo, a))]
    return np.asarray(c, np.float64)


def lpt_fields(initial_conditions, particles=None, order=1):
    """The force fields lpt scales: Psi1 = pm_forces(delta) read at the
    particles (pm.py:79-82) and, for order 2, Psi2 from the 2LPT source
    (pm.py:88-112).  Differentiable in the initial conditions."""
    ic = to_device(initial_conditions)
    shape = tuple(ic.shape)
    rel = particles is None
    parts = torch.zeros((*shape, 3), dtype=torch.float32, device=ic.device) if rel else to_device(particles)
    psi1 = _forces_from_field(ic, parts.reshape(-1, 3), rel, 0.0, 0).reshape(parts.shape)
    if order == 1:
        return psi1, None
    phi6 = _autograd.ShearField.apply(ic, 0) if _needs_grad(ic) else \
        _ops.c2r(_ops.kspace_shear6(_ops.r2c(ic), shape, 1.0 / ic.numel()), shape, clobber=True)
    delta2 = _autograd.Lpt2Source.apply(phi6) if _needs_grad(phi6) else _ops.lpt2_source(phi6)
    psi2 = _forces_from_field(delta2, parts.reshape(-1, 3), rel, 0.0, 0).reshape(parts.shape)
    return psi1, psi2


def lpt_combine(coeffs, psi1, psi2=None):
    """(dx, p, f) from the fields of :func:`lpt_fields` and a float64 host
    tensor of :func:`lpt_coefficients`, differentiable in both (the
    cotangents of the coefficients are float64 dot products)."""
    lc = _autograd.LinComb.apply
    one = torch.ones(2, dtype=torch.float64)
    dx = lc(coeffs[0:1], psi1)
    p = lc(coeffs[1:2], dx)
    f = lc(coeffs[2:3], psi1)
    if psi2 is None:
        return dx, p, f
    dx2 = lc(coeffs[3:4], psi2)
    p2 = lc(coeffs[4:5], dx2)
    f2 = lc(coeffs[5:6], psi2)
    return lc(one, dx, dx2), lc(one, p, p2), lc(one, f, f2)


class TabulatedPk:
    """``pk(x) = jnp.interp(x, k, pk)`` as a callable (the form every
    reference call site uses, e.g. tests/conftest.py:90-94).  ``linear_field``
    evaluates it inside the fused k-space kernel."""

    def __init__(self, k, pk):
        self.k = np.asarray(k, np.float64)
        self.pk = np.asarray(pk, np.float64)
        if self.k.shape != self.pk.shape or self.k.ndim != 1 or self.k.size < 2:
            raise ValueError("k and pk must be 1-D arrays of the same length (>= 2)")
        self._dev = None

    def __call__(self, x):
        if isinstance(x, torch.Tensor):
            return torch.as_tensor(np.interp(x.detach().cpu().numpy(), self.k, self.pk), device=x.device,
                                   dtype=x.dtype)
        return np.interp(x, self.k, self.pk)

    def device_tables(self, dev):
        if self._dev is None or self._dev[0].device != dev:
            self._dev = (torch.as_tensor(self.k, dtype=torch.float32, device=dev),
                         torch.as_tensor(self.pk, dtype=torch.float32, device=dev))
        return self._dev


class _SpectralMultiply(torch.autograd.Function):
    """x -> Re IFFT[FFT(x) * mult] for a real, even multiplier on the half
    grid (symmetric, so the backward applies the same map)."""

    @staticmethod
    def forward(ctx, x, mult):
        shape = tuple(x.shape)
        ctx.save_for_backward(mult)
        ctx.shape = shape
        return _ops.c2r(_ops.r2c(x.contiguous()) * mult, shape, clobber=True)

    @staticmethod
    def backward(ctx, g):
        (mult,) = ctx.saved_tensors
        return _ops.c2r(_ops.r2c(g.contiguous()) * mult, ctx.shape, clobber=True), None


def linear_field(mesh_shape, box_size, pk, seed, sharding=None):
    """pm.py:129-144: Gaussian initial conditions Re IFFT[FFT(w) sqrt(pk(|k|)
    M / V)].  ``seed`` is an int / torch.Generator (white noise from
    :func:`normal_field`) or the white-noise array itself, which lets a caller
    reproduce a JAX run exactly.  ``pk`` is a callable of |k| [h/Mpc] or a
    :class:`TabulatedPk` (evaluated inside the k-space kernel)."""
    single_device(sharding)
    shape = tuple(int(s) for s in mesh_shape)
    if isinstance(seed, (int, np.integer, torch.Generator)):
        noise = normal_field(seed, shape)
    else:
        noise = to_device(seed)
        if tuple(noise.shape) != shape:
            raise ValueError(f"noise shape {tuple(noise.shape)} does not match mesh_shape {shape}")
    box = tuple(float(b) for b in box_size)
    if isinstance(pk, TabulatedPk):
        ktab, pktab = pk.device_tables(noise.device)
        return _autograd.LinearField.apply(noise, box, ktab, pktab)
    # generic callable: evaluate pk on the half-grid |k| (pm.py:137-140)
    m = shape[0] * shape[1] * shape[2]
    vol = box[0] * box[1] * box[2]
    kv = []
    for d, n in enumerate(shape):
        f = torch.fft.fftfreq(n, dtype=torch.float64) if d < 2 else torch.fft.rfftfreq(n, dtype=torch.float64)
        k = (2 * np.pi * f).to(torch.float32)
        s = [1, 1, 1]
        s[d] = k.numel()
        kv.append((k / box[d] * n).reshape(s).to(noise.device))
    kmesh = (kv[0] ** 2 + kv[1] ** 2 + kv[2] ** 2) ** 0.5
    try:
        pkmesh = pk(kmesh)
    except Exception:
        pkmesh = pk(kmesh.cpu().numpy())
    pkmesh = to_device(pkmesh).reshape(kmesh.shape)
    mult = torch.sqrt(pkmesh * m / vol) * (1.0 / m)
    return _SpectralMultiply.apply(noise, mult)
