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
    def forward(ctx, meshes, pos, relative, deterministic, method="direct"):
        ctx.save_for_backward(meshes, pos)
        ctx.relative, ctx.deterministic = relative, deterministic
        return _ops.read(meshes, pos, relative=relative, method=method)

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
        return gm, gp, None, None, None


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
    tabulated pk.  The multiplier is real and even, so the map is symmetric in
    the noise; the cotangent of the pk table comes from
    jpm_kspace_linear_field_table_vjp_f32."""

    @staticmethod
    def forward(ctx, noise, box, ktab, pktab):
        shape = tuple(noise.shape)
        noise = noise.contiguous()
        fk = _ops.kspace_linear_field(_ops.r2c(noise), shape, box, ktab, pktab, 1.0 / noise.numel())
        ctx.save_for_backward(noise, ktab, pktab)
        ctx.box, ctx.shape = box, shape
        return _ops.c2r(fk, shape, clobber=True)

    @staticmethod
    def backward(ctx, g):
        noise, ktab, pktab = ctx.saved_tensors
        gk = _ops.r2c(g.contiguous())
        gn = gt = None
        if ctx.needs_input_grad[3]:
            gt = _ops.linear_field_table_vjp(_ops.r2c(noise), gk, ctx.shape, ctx.box, ktab, pktab,
                                             1.0 / g.numel()).to(pktab.dtype)
        if ctx.needs_input_grad[0]:
            fk = _ops.kspace_linear_field(gk, ctx.shape, ctx.box, ktab, pktab, 1.0 / g.numel())
            gn = _ops.c2r(fk, ctx.shape, clobber=True)
        return gn, None, None, gt


class LinComb(torch.autograd.Function):
    """y = sum_j c[j] x_j for float32 device tensors x_j of one shape and a
    host float64 coefficient vector c: the growth-factor products of lpt
    (pm.py:84-124).  Each term is added by jpm_axpy_f32 in order, so
    y = fl(fl(c0 x0) + fl(c1 x1)) as the reference's ``dx + dx2``; the
    cotangent of c[j] is <g, x_j> (jpm_dot_f32, float64)."""

    @staticmethod
    def forward(ctx, c, *xs):
        cv = [float(v) for v in c.detach().to(torch.float64).reshape(-1).tolist()]
        if len(cv) != len(xs):
            raise ValueError(f"{len(cv)} coefficients for {len(xs)} terms")
        xs = [x.contiguous() for x in xs]
        y = torch.zeros_like(xs[0])
        for cj, xj in zip(cv, xs):
            _ops.axpy(cj, xj, y)
        ctx.cv = cv
        ctx.c_meta = (c.dtype, c.device)
        ctx.save_for_backward(*xs)
        return y

    @staticmethod
    def backward(ctx, g):
        xs = ctx.saved_tensors
        g = g.contiguous()
        gc = None
        if ctx.needs_input_grad[0]:
            d = torch.empty(len(xs), dtype=torch.float64, device=g.device)
            for j, xj in enumerate(xs):
                _ops.dot_into(d, j, g, xj)
            gc = d.to(device=ctx.c_meta[1], dtype=ctx.c_meta[0])
        gxs = []
        for j, cj in enumerate(ctx.cv):
            if ctx.needs_input_grad[1 + j]:
                gxs.append(_ops.axpy(cj, g, torch.zeros_like(g)))
            else:
                gxs.append(None)
        return (gc, *gxs)


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


class CicPaint2d(torch.autograd.Function):
    """out = mesh + paint2d(pos, w) (painting.py:131-158)."""

    @staticmethod
    def forward(ctx, mesh, pos2, weight):
        out = _ops.paint_2d(mesh.clone(), pos2, weight)
        ctx.save_for_backward(pos2, weight)
        return out

    @staticmethod
    def backward(ctx, g):
        pos2, weight = ctx.saved_tensors
        g = g.contiguous()
        gpos, gw = _ops.paint_2d_vjp(g, pos2, weight, want_pos=ctx.needs_input_grad[1],
                                     want_weight=ctx.needs_input_grad[2] and weight is not None)
        return (g if ctx.needs_input_grad[0] else None), gpos, gw


class DensityPlane(torch.autograd.Function):
    """The unsmoothed density plane of lensing.py:11-40."""

    @staticmethod
    def forward(ctx, pos, box_shape, center, width, res):
        ctx.save_for_backward(pos)
        ctx.args = (box_shape, center, width, res)
        return _ops.density_plane(pos, box_shape, center, width, res)

    @staticmethod
    def backward(ctx, g):
        (pos,) = ctx.saved_tensors
        gpos = _ops.density_plane_vjp(g.contiguous(), pos, *ctx.args) if ctx.needs_input_grad[0] else None
        return gpos, None, None, None, None


class GaussianSmoothing(torch.autograd.Function):
    """utils.py:208-222; the filter is real and even, so the backward is the
    same filter applied to the cotangent."""

    @staticmethod
    def forward(ctx, im, sigma):
        ctx.sigma = sigma
        return _ops.gaussian_smoothing(im, sigma)

    @staticmethod
    def backward(ctx, g):
        return _ops.gaussian_smoothing(g.contiguous(), ctx.sigma), None
