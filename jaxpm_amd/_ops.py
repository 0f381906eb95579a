"""Thin tensor-level wrappers over the C ABI (one function per entry point).

Every wrapper takes contiguous float32 / complex64 device tensors, allocates
its outputs with torch on the same device and launches on torch's current
stream.  Shapes are validated here and again by the library.
"""
import ctypes

import torch

from . import _lib
from ._lib import call, ptr, size, stream


def half_shape(shape):
    nx, ny, nz = shape
    return (nx, ny, nz // 2 + 1)


def _f32(t, name):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a device tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    return t.contiguous()


def _c64(t, name):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a device tensor")
    if t.dtype != torch.complex64:
        raise TypeError(f"{name} must be complex64, got {t.dtype}")
    return t.contiguous()


def _dims(shape):
    if len(shape) != 3:
        raise ValueError(f"mesh shape must have 3 dimensions, got {tuple(shape)}")
    return [int(s) for s in shape]


# ---------------------------------------------------------------- painting

def binned_ok(shape):
    """The binned paint indexes three meshes with 32-bit offsets."""
    m = 1
    for s in shape:
        m *= int(s)
    return 3 * m <= 2 ** 32


def paint(mesh, pos, weight=None, wscalar=1.0, relative=False, deterministic=False, zero=False, method="binned"):
    """mesh (nx,ny,nz) f32 updated in place with CIC deposits of pos (n,3).

    method "binned" (tile sort + LDS windows, the default) or "atomic"
    (one global atomic per corner)."""
    mesh = _f32(mesh, "mesh")
    pos = _f32(pos, "positions").reshape(-1, 3)
    nx, ny, nz = _dims(mesh.shape)
    n = pos.shape[0]
    if weight is not None:
        weight = _f32(weight, "weight").reshape(-1)
        if weight.numel() != n:
            raise ValueError(f"weight has {weight.numel()} entries for {n} particles")
    flags = (_lib.RELATIVE if relative else 0) | (_lib.DETERMINISTIC if deterministic else 0) | \
        (_lib.ZERO_MESH if zero else 0)
    if method == "binned" and binned_ok(mesh.shape):
        wsb = size("jpm_cic_paint_binned_workspace_size", nx, ny, nz, n, 1, flags)
        ws = torch.empty(wsb, dtype=torch.uint8, device=mesh.device)
        call("jpm_cic_paint_binned_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(weight), float(wscalar), 1,
             ptr(mesh), flags, ptr(ws), wsb)
        return mesh
    if method not in ("binned", "atomic"):
        raise ValueError(f"unknown paint method {method!r}")
    wsb = size("jpm_cic_paint_workspace_size", nx, ny, nz, n, flags)
    ws = torch.empty(wsb, dtype=torch.uint8, device=mesh.device) if wsb else None
    call("jpm_cic_paint_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(weight), float(wscalar), ptr(mesh),
         flags, ptr(ws), wsb)
    return mesh


def paint_f64(mesh, pos, weight=None, wscalar=1.0, relative=False, zero=False):
    """Float64 paint into ``mesh`` (float64, updated in place)."""
    nx, ny, nz = _dims(mesh.shape)
    n = pos.shape[0]
    flags = (_lib.RELATIVE if relative else 0) | (_lib.ZERO_MESH if zero else 0)
    call("jpm_cic_paint_f64", stream(), nx, ny, nz, ptr(pos), n, ptr(weight), float(wscalar), ptr(mesh), flags)
    return mesh


def read_f64(mesh4, pos, relative=False):
    """Float64 read of nc meshes (nc, nx, ny, nz) at pos (n, 3) -> (n, nc)."""
    nc = mesh4.shape[0]
    nx, ny, nz = _dims(mesh4.shape[1:])
    n = pos.shape[0]
    out = torch.empty((n, nc), dtype=torch.float64, device=pos.device)
    call("jpm_cic_read_f64", stream(), nx, ny, nz, ptr(mesh4), nc, nx * ny * nz, ptr(pos), n, ptr(out),
         _lib.RELATIVE if relative else 0)
    return out


def pm_forces_f64(pos, shape, delta=None, r_split=0.0, relative=False, kf=0):
    """Float64 pm_forces: (n, 3) forces at pos from its paint (or the real
    density ``delta``)."""
    nx, ny, nz = _dims(shape)
    n = pos.shape[0]
    wsb = size("jpm_pm_forces_f64_workspace_size", nx, ny, nz)
    ws = torch.empty(wsb, dtype=torch.uint8, device=pos.device)
    out = torch.empty((n, 3), dtype=torch.float64, device=pos.device)
    call("jpm_pm_forces_f64", stream(), nx, ny, nz, ptr(pos), n, ptr(delta), float(r_split),
         _lib.RELATIVE if relative else 0, kf, ptr(out), ptr(ws), wsb)
    return out


def paint_multi(shape, pos, weights, relative=False, deterministic=False, method="binned"):
    """nc meshes (nc,nx,ny,nz): channel c painted with weights[:, c]."""
    nx, ny, nz = _dims(shape)
    pos = _f32(pos, "positions").reshape(-1, 3)
    weights = _f32(weights, "weights")
    n = pos.shape[0]
    weights = weights.reshape(n, -1)
    nc = weights.shape[1]
    out = torch.empty((nc, nx, ny, nz), dtype=torch.float32, device=pos.device)
    flags = (_lib.RELATIVE if relative else 0) | (_lib.DETERMINISTIC if deterministic else 0) | _lib.ZERO_MESH
    if method == "binned" and nc in (1, 3) and binned_ok(shape):
        wsb = size("jpm_cic_paint_binned_workspace_size", nx, ny, nz, n, nc, flags)
        ws = torch.empty(wsb, dtype=torch.uint8, device=pos.device)
        call("jpm_cic_paint_binned_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(weights), 1.0, nc, ptr(out),
             flags, ptr(ws), wsb)
        return out
    wsb = size("jpm_cic_paint_multi_workspace_size", nx, ny, nz, nc, flags)
    ws = torch.empty(wsb, dtype=torch.uint8, device=pos.device) if wsb else None
    call("jpm_cic_paint_multi_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(weights), nc, ptr(out), flags,
         ptr(ws), wsb)
    return out


def read(meshes, pos, relative=False, method="direct"):
    """meshes (nc,nx,ny,nz) or (nx,ny,nz) -> values (n,nc) or (n,).

    method "direct" (8 corner loads per particle, the default) or "sorted"
    (a tile sort first, then LDS windows: faster for particles in no spatial
    order; absolute positions, up to 3 meshes; the same values)."""
    meshes = _f32(meshes, "mesh")
    single = meshes.dim() == 3
    m4 = meshes.unsqueeze(0) if single else meshes
    nc = m4.shape[0]
    nx, ny, nz = _dims(m4.shape[1:])
    pos = _f32(pos, "positions").reshape(-1, 3)
    n = pos.shape[0]
    out = torch.empty((n, nc), dtype=torch.float32, device=pos.device)
    if method == "sorted":
        if relative or nc > 3:
            raise ValueError("the sorted read takes absolute positions and at most 3 meshes")
        wsb = size("jpm_cic_read_sorted_workspace_size", nx, ny, nz, n)
        if wsb == 0 and n > 0:
            raise ValueError(f"mesh {(nx, ny, nz)} is too large for the tile-sorted read")
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=pos.device)
        call("jpm_cic_read_sorted_f32", stream(), nx, ny, nz, ptr(m4), nc, nx * ny * nz, ptr(pos), n, ptr(out), 0,
             ptr(ws), wsb)
        return out.reshape(n) if single else out
    if method != "direct":
        raise ValueError(f"unknown read method {method!r}")
    call("jpm_cic_read_f32", stream(), nx, ny, nz, ptr(m4), nc, nx * ny * nz, ptr(pos), n, ptr(out),
         _lib.RELATIVE if relative else 0)
    return out.reshape(n) if single else out


def grad_pos(meshes, pos, cot=None, weight=None, wscalar=1.0, relative=False, out=None):
    """Position cotangent (n,3) of CIC sampling (see jpm_cic_grad_pos_f32)."""
    meshes = _f32(meshes, "mesh")
    m4 = meshes.unsqueeze(0) if meshes.dim() == 3 else meshes
    nc = m4.shape[0]
    nx, ny, nz = _dims(m4.shape[1:])
    pos = _f32(pos, "positions").reshape(-1, 3)
    n = pos.shape[0]
    if cot is not None:
        cot = _f32(cot, "cotangent").reshape(n, nc)
    if weight is not None:
        weight = _f32(weight, "weight").reshape(n)
    acc = out is not None
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float32, device=pos.device)
    call("jpm_cic_grad_pos_f32", stream(), nx, ny, nz, ptr(m4), nc, nx * ny * nz, ptr(pos), n, ptr(cot),
         ptr(weight), float(wscalar), ptr(out), _lib.RELATIVE if relative else 0, int(acc))
    return out


# --------------------------------------------------------------------- FFT

def r2c(x):
    """(…, nx, ny, nz) f32 -> (…, nx, ny, nz//2+1) c64, unnormalised."""
    x = _f32(x, "field")
    nx, ny, nz = _dims(x.shape[-3:])
    batch = x.numel() // (nx * ny * nz)
    out = torch.empty((*x.shape[:-3], *half_shape((nx, ny, nz))), dtype=torch.complex64, device=x.device)
    call("jpm_fft_r2c_f32", stream(), nx, ny, nz, batch, ptr(x), ptr(out))
    return out


def c2r(xk, shape, clobber=False):
    """Half spectra -> real meshes of `shape` (unnormalised).  The input is
    copied first unless ``clobber`` (the transform overwrites its input)."""
    xk = _c64(xk, "spectrum")
    nx, ny, nz = _dims(shape)
    if tuple(xk.shape[-3:]) != half_shape((nx, ny, nz)):
        raise ValueError(f"spectrum shape {tuple(xk.shape)} does not match mesh {tuple(shape)}")
    if not clobber:
        xk = xk.clone()
    batch = xk.numel() // xk.shape[-3:].numel()
    out = torch.empty((*xk.shape[:-3], nx, ny, nz), dtype=torch.float32, device=xk.device)
    call("jpm_fft_c2r_f32", stream(), nx, ny, nz, batch, ptr(xk), ptr(out))
    return out


def c2c(xk, inverse=False):
    """Unnormalised complex-to-complex transform of (…, nx, ny, nz) c64."""
    xk = _c64(xk, "spectrum")
    nx, ny, nz = _dims(xk.shape[-3:])
    batch = xk.numel() // (nx * ny * nz)
    out = torch.empty_like(xk)
    ws = torch.empty(size("jpm_fft_c2c_workspace_size", nx, ny, nz, batch), dtype=torch.uint8, device=xk.device)
    call("jpm_fft_c2c_f32", stream(), nx, ny, nz, batch, ptr(xk), ptr(out), 1 if inverse else -1, ptr(ws), ws.numel())
    return out


def full_to_half(xk):
    xk = _c64(xk, "spectrum")
    nx, ny, nz = _dims(xk.shape)
    out = torch.empty(half_shape((nx, ny, nz)), dtype=torch.complex64, device=xk.device)
    call("jpm_spectrum_full_to_half_f32", stream(), nx, ny, nz, ptr(xk), ptr(out))
    return out


def half_to_full(hk, shape):
    hk = _c64(hk, "spectrum")
    nx, ny, nz = _dims(shape)
    out = torch.empty((nx, ny, nz), dtype=torch.complex64, device=hk.device)
    call("jpm_spectrum_half_to_full_f32", stream(), nx, ny, nz, ptr(hk), ptr(out))
    return out


# ----------------------------------------------------------------- k-space

def kflags(order=1, fd=False, deconv=False):
    return (_lib.K_GRAD_EXACT if order == 0 else 0) | (_lib.K_LAPL_FD if fd else 0) | \
        (_lib.K_DECONV if deconv else 0)


def kspace_force(dk, shape, r_split=0.0, scale=1.0, flags=0):
    dk = _c64(dk, "delta_k")
    nx, ny, nz = _dims(shape)
    out = torch.empty((3, *half_shape(shape)), dtype=torch.complex64, device=dk.device)
    call("jpm_kspace_force_f32", stream(), nx, ny, nz, ptr(dk), ptr(out), float(r_split), float(scale), flags)
    return out


def kspace_potential(dk, shape, r_split=0.0, scale=1.0, flags=0):
    dk = _c64(dk, "delta_k")
    nx, ny, nz = _dims(shape)
    out = torch.empty(half_shape(shape), dtype=torch.complex64, device=dk.device)
    call("jpm_kspace_potential_f32", stream(), nx, ny, nz, ptr(dk), ptr(out), float(r_split), float(scale), flags)
    return out


def fd4_gradient(phi, scale=1.0):
    """(nx,ny,nz) -> (3,nx,ny,nz): -scale * 4-point central difference."""
    phi = _f32(phi, "potential")
    nx, ny, nz = _dims(phi.shape)
    out = torch.empty((3, nx, ny, nz), dtype=torch.float32, device=phi.device)
    call("jpm_fd4_gradient_f32", stream(), nx, ny, nz, ptr(phi), float(scale), ptr(out))
    return out


def fd4_divergence(g3, scale=1.0):
    g3 = _f32(g3, "meshes")
    nx, ny, nz = _dims(g3.shape[1:])
    out = torch.empty((nx, ny, nz), dtype=torch.float32, device=g3.device)
    call("jpm_fd4_divergence_f32", stream(), nx, ny, nz, ptr(g3), float(scale), ptr(out))
    return out


def kspace_force_adjoint(in3, shape, r_split=0.0, scale=1.0, flags=0):
    in3 = _c64(in3, "spectra")
    nx, ny, nz = _dims(shape)
    out = torch.empty(half_shape(shape), dtype=torch.complex64, device=in3.device)
    call("jpm_kspace_force_adjoint_f32", stream(), nx, ny, nz, ptr(in3), ptr(out), float(r_split), float(scale),
         flags)
    return out


def kspace_compensate(fk, shape, scale=1.0):
    fk = _c64(fk, "spectrum")
    nx, ny, nz = _dims(shape)
    call("jpm_kspace_compensate_f32", stream(), nx, ny, nz, ptr(fk), float(scale))
    return fk


def kspace_linear_field(fk, shape, box, ktab, pktab, scale=1.0):
    fk = _c64(fk, "spectrum")
    nx, ny, nz = _dims(shape)
    ktab = _f32(ktab, "k table")
    pktab = _f32(pktab, "pk table")
    if ktab.numel() != pktab.numel():
        raise ValueError("k and P(k) tables differ in length")
    call("jpm_kspace_linear_field_f32", stream(), nx, ny, nz, ptr(fk), float(box[0]), float(box[1]), float(box[2]),
         ptr(ktab), ptr(pktab), ktab.numel(), float(scale))
    return fk


def kspace_shear6(dk, shape, scale=1.0, flags=0):
    dk = _c64(dk, "delta_k")
    nx, ny, nz = _dims(shape)
    out = torch.empty((6, *half_shape(shape)), dtype=torch.complex64, device=dk.device)
    call("jpm_kspace_shear6_f32", stream(), nx, ny, nz, ptr(dk), ptr(out), float(scale), flags)
    return out


def kspace_shear6_adjoint(in6, shape, scale=1.0, flags=0):
    in6 = _c64(in6, "spectra")
    nx, ny, nz = _dims(shape)
    out = torch.empty(half_shape(shape), dtype=torch.complex64, device=in6.device)
    call("jpm_kspace_shear6_adjoint_f32", stream(), nx, ny, nz, ptr(in6), ptr(out), float(scale), flags)
    return out


def lpt2_source(phi6):
    phi6 = _f32(phi6, "shear meshes")
    m = phi6[0].numel()
    out = torch.empty(phi6.shape[1:], dtype=torch.float32, device=phi6.device)
    call("jpm_lpt2_source_f32", stream(), m, ptr(phi6), ptr(out))
    return out


def lpt2_source_vjp(phi6, gbar):
    phi6 = _f32(phi6, "shear meshes")
    gbar = _f32(gbar, "cotangent")
    m = phi6[0].numel()
    out = torch.empty_like(phi6)
    call("jpm_lpt2_source_vjp_f32", stream(), m, ptr(phi6), ptr(gbar), ptr(out))
    return out


# ---------------------------------------------------------------- pipeline

def pm_forces_fused(pos, shape, delta_k=None, r_split=0.0, scale=1.0, relative=False, deterministic=False,
                    kf=0, out=None, ws=None, atomic=False):
    """forces (n,3) through the single-call pipeline jpm_pm_forces_f32."""
    pos = _f32(pos, "positions").reshape(-1, 3)
    nx, ny, nz = _dims(shape)
    n = pos.shape[0]
    pflags = (_lib.RELATIVE if relative else 0) | (_lib.DETERMINISTIC if deterministic else 0) | \
        (_lib.ATOMIC if atomic else 0)
    flags = pflags | (kf << 8)
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float32, device=pos.device)
    if delta_k is None and n > 0:
        # tile-sorted pipeline: potential + LDS stencil gather (writes forces)
        kb = size("jpm_force_kick_workspace_size", nx, ny, nz, n, flags)
        if kb:
            if ws is None or ws.numel() < kb:
                ws = torch.empty(kb, dtype=torch.uint8, device=pos.device)
            phi = torch.empty((nx, ny, nz), dtype=torch.float32, device=pos.device)
            call("jpm_force_kick_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(out), float(scale), ptr(phi),
                 float(r_split), flags, 1, ptr(ws), ws.numel(), None, 0, None)
            return out
    wsb = size("jpm_pm_forces_workspace_size", nx, ny, nz, n, flags)
    if ws is None or ws.numel() < wsb:
        ws = torch.empty(wsb, dtype=torch.uint8, device=pos.device)
    if delta_k is not None:
        delta_k = _c64(delta_k, "delta_k")
    call("jpm_pm_forces_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(delta_k), ptr(out), float(r_split),
         float(scale), flags, ptr(ws), ws.numel())
    return out


# ---------------------------------------------------------------- analysis

def power_spectrum_sums(spec1, spec2, shape, fac, edges, poles, los):
    """Per-bin mode sums over half spectra (jpm_power_spectrum_f32).

    Returns a float64 device tensor [slot][bin]: count, sum |k|, Re sums per
    pole and, for a cross spectrum, Im sums per pole."""
    spec1 = _c64(spec1, "spectrum")
    nx, ny, nz = _dims(shape)
    if tuple(spec1.shape) != half_shape((nx, ny, nz)):
        raise ValueError(f"spectrum shape {tuple(spec1.shape)} does not match mesh {tuple(shape)}")
    if spec2 is not None:
        spec2 = _c64(spec2, "spectrum2")
        if spec2.shape != spec1.shape:
            raise ValueError("the two spectra differ in shape")
    edges = torch.as_tensor(edges, dtype=torch.float64).to(spec1.device).contiguous()
    ne = int(edges.numel())
    npoles = len(poles)
    cross = int(spec2 is not None)
    slots = 2 + npoles * (2 if cross else 1)
    out = torch.empty((slots, ne + 1), dtype=torch.float64, device=spec1.device)
    ws = torch.empty(size("jpm_power_spectrum_workspace_size", nx, ny, nz, ne, npoles, cross),
                     dtype=torch.uint8, device=spec1.device)
    parr = (ctypes.c_int * max(npoles, 1))(*[int(p) for p in poles])
    call("jpm_power_spectrum_f32", stream(), nx, ny, nz, ptr(spec1), ptr(spec2) if cross else None,
         float(fac[0]), float(fac[1]), float(fac[2]), ptr(edges) if ne else None, ne, parr, npoles,
         float(los[0]), float(los[1]), float(los[2]), ptr(out), ptr(ws), ws.numel())
    return out


def linear_field_table_vjp(noise_k, cot_k, shape, box, ktab, pktab, scale=1.0):
    """float64 (ntab,) cotangent of the pk table of kspace_linear_field + C2R
    (jpm_kspace_linear_field_table_vjp_f32)."""
    noise_k = _c64(noise_k, "noise spectrum")
    cot_k = _c64(cot_k, "cotangent spectrum")
    nx, ny, nz = _dims(shape)
    ktab = _f32(ktab, "k table")
    pktab = _f32(pktab, "pk table")
    nt = ktab.numel()
    if pktab.numel() != nt:
        raise ValueError("k and P(k) tables differ in length")
    out = torch.empty(nt, dtype=torch.float64, device=noise_k.device)
    wsb = size("jpm_linear_field_table_vjp_workspace_size", nt)
    ws = torch.empty(wsb, dtype=torch.uint8, device=noise_k.device)
    call("jpm_kspace_linear_field_table_vjp_f32", stream(), nx, ny, nz, ptr(noise_k), ptr(cot_k), float(box[0]),
         float(box[1]), float(box[2]), ptr(ktab), ptr(pktab), nt, float(scale), ptr(out), ptr(ws), wsb)
    return out


_dot_ws = {}


def dot_into(out, i, x, y):
    """out[i] (a float64 device tensor) = <x, y> over float32 tensors of one
    size, accumulated in float64 (jpm_dot_f32)."""
    x = _f32(x, "x")
    y = _f32(y, "y")
    if x.numel() != y.numel():
        raise ValueError(f"dot of {x.numel()} and {y.numel()} elements")
    if out.dtype != torch.float64 or not out.is_contiguous():
        raise TypeError("dot output must be a contiguous float64 tensor")
    ws = _dot_ws.get(x.device)
    if ws is None:
        ws = _dot_ws[x.device] = torch.empty(size("jpm_dot_workspace_size"), dtype=torch.uint8, device=x.device)
    call("jpm_dot_f32", stream(), x.numel(), ptr(x), ptr(y), ctypes.c_void_p(out.data_ptr() + 8 * int(i)), ptr(ws),
         ws.numel())


def paint_2d(mesh, pos2, weight=None):
    """mesh[nx][ny] += the 2-D CIC deposits of pos2 (n, 2), times weight (n)
    when given, in place (jpm_cic_paint_2d_f32)."""
    if mesh.dim() != 2 or mesh.dtype != torch.float32 or not mesh.is_contiguous():
        raise ValueError("the plane must be a contiguous 2-D float32 tensor")
    pos2 = _f32(pos2, "positions")
    n = pos2.numel() // 2
    w = None if weight is None else _f32(weight, "weight")
    if w is not None and w.numel() != n:
        raise ValueError("Weight shape must match particle shape")
    call("jpm_cic_paint_2d_f32", stream(), mesh.shape[0], mesh.shape[1], ptr(pos2), n, 2, ptr(w), ptr(mesh))
    return mesh


def paint_2d_vjp(g, pos2, weight=None, want_pos=True, want_weight=False):
    """(gpos (n, 2) or None, gweight (n) or None) of the 2-D paint for the
    plane cotangent g (jpm_cic_paint_2d_vjp_f32)."""
    g = _f32(g, "cotangent")
    pos2 = _f32(pos2, "positions")
    n = pos2.numel() // 2
    w = None if weight is None else _f32(weight, "weight")
    gpos = torch.empty((n, 2), dtype=torch.float32, device=g.device) if want_pos else None
    gw = torch.empty(n, dtype=torch.float32, device=g.device) if want_weight else None
    if want_pos or want_weight:
        call("jpm_cic_paint_2d_vjp_f32", stream(), g.shape[0], g.shape[1], ptr(pos2), n, 2, ptr(w), ptr(g),
             ptr(gpos), 2, ptr(gw))
    return gpos, gw


def density_plane(pos, box_shape, center, width, res):
    """A new res x res plane (jpm_density_plane_f32) of positions (n, 3)."""
    pos = _f32(pos, "positions")
    plane = torch.empty((res, res), dtype=torch.float32, device=pos.device)
    call("jpm_density_plane_f32", stream(), ptr(pos), pos.numel() // 3, int(box_shape[0]), int(box_shape[1]),
         float(center), float(width), int(res), ptr(plane))
    return plane


def density_plane_vjp(g, pos, box_shape, center, width, res):
    """The position cotangent (n, 3) of density_plane for the plane
    cotangent g (jpm_density_plane_vjp_f32; g is copied first)."""
    pos = _f32(pos, "positions")
    g = _f32(g, "cotangent").clone()
    gpos = torch.empty_like(pos)
    call("jpm_density_plane_vjp_f32", stream(), ptr(pos), pos.numel() // 3, int(box_shape[0]), int(box_shape[1]),
         float(center), float(width), int(res), ptr(g), ptr(gpos))
    return gpos


def gaussian_smoothing(im, sigma):
    """A new image: the Gaussian filter of utils.gaussian_smoothing applied
    to the 2-D float32 image im (jpm_gaussian_smoothing_f32)."""
    im = _f32(im, "image")
    if im.dim() != 2:
        raise ValueError("gaussian_smoothing takes a 2-D image")
    n0, n1 = im.shape
    ws = torch.empty(size("jpm_gaussian_smoothing_workspace_size", n0, n1), dtype=torch.uint8, device=im.device)
    out = torch.empty_like(im)
    call("jpm_gaussian_smoothing_f32", stream(), n0, n1, ptr(im), ptr(out), float(sigma), ptr(ws), ws.numel())
    return out


def tile_order(pos, mesh_shape, ws=None):
    """int32 tensor (uint32 bits) order[k] = index of the particle at slot k
    of the tile sort of absolute positions ``pos`` (jpm_tile_order_u32).
    ``ws``: an optional uint8 workspace of at least
    jpm_tile_order_workspace_size bytes."""
    pos = _f32(pos, "positions")
    nx, ny, nz = _dims(mesh_shape)
    n = pos.numel() // 3
    need = size("jpm_tile_order_workspace_size", nx, ny, nz, n)
    if need == 0:
        raise ValueError(f"mesh {tuple(mesh_shape)} too large for the tile sort")
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=pos.device)
    order = torch.empty(n, dtype=torch.int32, device=pos.device)
    call("jpm_tile_order_u32", stream(), nx, ny, nz, ptr(pos), n, ptr(order), ptr(ws), ws.numel())
    return order


def permute_rows(order, *arrays, scatter=False, out=None):
    """(n, 3) float32 tensors, one per array: out[k] = x[order[k]], or with
    ``scatter`` out[order[k]] = x[k] (the inverse); one or two arrays per
    launch (jpm_permute_rows_f32).  ``out``: the output tensors (new ones
    when None)."""
    if order.dtype != torch.int32 or not order.is_contiguous():
        raise ValueError("order must be a contiguous int32 tensor")
    xs = [_f32(x, "rows") for x in arrays]
    if any(3 * order.numel() != x.numel() for x in xs):
        raise ValueError("order must have one entry per row")
    if out is None:
        outs = [torch.empty_like(x) for x in xs]
    else:
        outs = list(out)
        if len(outs) != len(xs) or any(o.dtype != torch.float32 or not o.is_contiguous() or o.numel() != x.numel()
                                       for o, x in zip(outs, xs)):
            raise ValueError("one contiguous float32 output of the input's size per array")
    for i in range(0, len(xs), 2):
        a, b = xs[i:i + 2], outs[i:i + 2]
        call("jpm_permute_rows_f32", stream(), order.numel(), ptr(order), ptr(a[0]), ptr(b[0]),
             ptr(a[1]) if len(a) > 1 else None, ptr(b[1]) if len(b) > 1 else None, 1 if scatter else 0)
    return outs[0] if len(outs) == 1 else tuple(outs)


def axpy(a, x, y):
    """y += a x in place (jpm_axpy_f32); float32 tensors of one size."""
    x = _f32(x, "x")
    if not y.is_contiguous() or y.dtype != torch.float32 or y.numel() != x.numel():
        raise ValueError("axpy target must be a contiguous float32 tensor of the same size")
    call("jpm_axpy_f32", stream(), x.numel(), float(a), ptr(x), ptr(y))
    return y
