"""Hand-written adjoints of the PM building blocks (torch.autograd.Function).

These play the role ``jax.custom_vjp`` would in a JAX binding.  Paint and
read are each other's transposes; the Fourier operators reuse the forward
FFT plans with the conjugate multiplier.  No backward pass here calls torch's
own autograd on device arrays -- every cotangent is produced by a kernel of
libjaxpm_hip.so.
"""
import torch

from . import _ops


class CicPaint(torch.autograd.Function):
    """out = mesh + paint(pos, w) (painting.py:15-45; relative: 161-189)."""

    @staticmethod
    def forward(ctx, mesh, pos, weight, wscalar, relative, deterministic):
        out = mesh.clone()
        _ops.paint(out, pos, weight=weight, wscalar=wscalar, relative=relative, deterministic=deterministic)
        ctx.save_for_backward(pos, weight)
        ctx.wscalar, ctx.relative, ctx.shape = wscalar, relative, pos.shape
        return out

    @staticmethod
    def backward(ctx, g):
        pos, weight = ctx.saved_tensors
        g = g.contiguous()
        gmesh = g if ctx.needs_input_grad[0] else None
        gpos = gw = None
        if ctx.needs_input_grad[1]:
            gpos = _ops.grad_pos(g, pos, weight=weight, wscalar=ctx.wscalar,
                                 relative=ctx.relative).reshape(ctx.shape)
        if ctx.needs_input_grad[2]:
            gw = _ops.read(g, pos, relative=ctx.relative).reshape(weight.shape)
        return gmesh, gpos, gw, None, None, None


class CicRead(torch.autograd.Function):
    """vals[p, c] = sum_o mesh_c[idx_o] W_o (painting.py:78-106; relative:
    218-236).  meshes: (nc, nx, ny, nz)."""

    @staticmethod
    def forward(ctx, meshes, pos, relative, deterministic):
        ctx.save_for_backward(meshes, pos)
        ctx.relative, ctx.deterministic = relative, deterministic
        return _ops.read(meshes, pos, relative=relative)

    @staticmethod
    def backward(ctx, g):
        meshes, pos = ctx.saved_tensors
        g = g.contiguous()
        gm = gp = None
        if ctx.needs_input_grad[0]:
            gm = _ops.paint_multi(meshes.shape[1:], pos, g, relative=ctx.relative,
                                  deterministic=ctx.deterministic)
        if ctx.needs_input_grad[1]:
            gp = _ops.grad_pos(meshes, pos, cot=g, relative=ctx.relative).reshape(pos.shape)
        return gm, gp, None, None


class ForceField(torch.autograd.Function):
    """rho (nx,ny,nz) -> phi (3,nx,ny,nz), phi_d = Re IFFT[-g_d G Lr FFT rho]
    (pm.py:49-56).  Linear; its transpose uses the conjugate multiplier."""

    @staticmethod
    def forward(ctx, rho, r_split, kf):
        shape = tuple(rho.shape)
        m = rho.numel()
        spec3 = _ops.kspace_force(_ops.r2c(rho.contiguous()), shape, r_split, 1.0 / m, kf)
        ctx.r_split, ctx.kf, ctx.shape = r_split, kf, shape
        return _ops.c2r(spec3, shape, clobber=True)

    @staticmethod
    def backward(ctx, g):
        m = g[0].numel()
        gk = _ops.kspace_force_adjoint(_ops.r2c(g.contiguous()), ctx.shape, ctx.r_split, 1.0 / m, ctx.kf)
        return _ops.c2r(gk, ctx.shape, clobber=True), None, None


class ShearField(torch.autograd.Function):
    """field -> six meshes Re IFFT[g_a g_b G FFT field] (pm.py:89-109)."""

    @staticmethod
    def forward(ctx, field, kf):
        shape = tuple(field.shape)
        spec6 = _ops.kspace_shear6(_ops.r2c(field.contiguous()), shape, 1.0 / field.numel(), kf)
        ctx.kf, ctx.shape = kf, shape
        return _ops.c2r(spec6, shape, clobber=True)

    @staticmethod
    def backward(ctx, g):
        m = g[0].numel()
        gk = _ops.kspace_shear6_adjoint(_ops.r2c(g.contiguous()), ctx.shape, 1.0 / m, ctx.kf)
        return _ops.c2r(gk, ctx.shape, clobber=True), None


class Lpt2Source(torch.autograd.Function):
    """six shear meshes -> delta2 (pm.py:92-109)."""

    @staticmethod
    def forward(ctx, phi6):
        ctx.save_for_backward(phi6)
        return _ops.lpt2_source(phi6)

    @staticmethod
    def backward(ctx, g):
        (phi6,) = ctx.saved_tensors
        return _ops.lpt2_source_vjp(phi6, g.contiguous())


class LinearField(torch.autograd.Function):
    """noise -> Re IFFT[FFT(noise) sqrt(pk(|k|) M / V)] (pm.py:129-144) with a
    tabulated pk.  The multiplier is real and even, so the map is symmetric."""

    @staticmethod
    def forward(ctx, noise, box, ktab, pktab):
        shape = tuple(noise.shape)
        fk = _ops.kspace_linear_field(_ops.r2c(noise.contiguous()), shape, box, ktab, pktab,
                                      1.0 / noise.numel())
        ctx.save_for_backward(ktab, pktab)
        ctx.box, ctx.shape = box, shape
        return _ops.c2r(fk, shape, clobber=True)

    @staticmethod
    def backward(ctx, g):
        ktab, pktab = ctx.saved_tensors
        fk = _ops.kspace_linear_field(_ops.r2c(g.contiguous()), ctx.shape, ctx.box, ktab, pktab,
                                      1.0 / g.numel())
        return _ops.c2r(fk, ctx.shape, clobber=True), None, None, None


class Compensate(torch.autograd.Function):
    """compensate_cic (painting.py:263-275); symmetric real-even multiplier."""

    @staticmethod
    def forward(ctx, field):
        shape = tuple(field.shape)
        ctx.shape = shape
        fk = _ops.kspace_compensate(_ops.r2c(field.contiguous()), shape, 1.0 / field.numel())
        return _ops.c2r(fk, shape, clobber=True)

    @staticmethod
    def backward(ctx, g):
        fk = _ops.kspace_compensate(_ops.r2c(g.contiguous()), ctx.shape, 1.0 / g.numel())
        return _ops.c2r(fk, ctx.shape, clobber=True)
