"""CIC painting and reading, mirroring ``jaxpm/painting.py``.

Same names, argument meaning and error behaviour as the reference; arrays
are torch device tensors (NumPy inputs are copied to the GPU).  The
reference's functional semantics hold: inputs are never modified, results
are fresh tensors.  Every call runs HIP kernels of libjaxpm_hip.so; gradients
flow through hand-written adjoints (``_autograd``).

Additive keywords: ``deterministic=True`` selects the bitwise-reproducible
fixed-point paint.  Particle counts other than the mesh size are accepted by
the absolute painter (the reference reshapes positions to the mesh shape,
painting.py:57, and so requires N == M).

With ``jaxpm_amd.config.update("jax_enable_x64", True)``, float64 input to
cic_paint / cic_read / cic_paint_dx / cic_read_dx runs the float64 kernels
and returns float64 (forward only), as the reference does under x64.
"""
import torch

from . import _autograd, _ops
from ._convert import forward_only, is_scalar, scalar_value, single_device, to_device, wants_f64


def _weight(weight, npart, dtype=torch.float32):
    if is_scalar(weight):
        return None, scalar_value(weight)
    w = to_device(weight, dtype).reshape(-1)
    if w.numel() != npart:
        raise ValueError("Weight shape must match particle shape")
    return w, 1.0


def cic_paint(grid_mesh, positions, weight=1., halo_size=0, sharding=None, deterministic=False):
    """Absolute CIC paint (painting.py:48-75 / 15-45): ``grid_mesh`` plus the
    deposits of ``positions`` (..., 3) in grid units, periodic wrap."""
    single_device(sharding)
    f64 = wants_f64(grid_mesh, positions, weight)
    dt = torch.float64 if f64 else torch.float32
    mesh = to_device(grid_mesh, dt)
    if mesh.dim() != 3:
        raise ValueError(f"grid_mesh must be 3-D, got shape {tuple(mesh.shape)}")
    pos = to_device(positions, dt)
    if pos.shape[-1] != 3:
        raise ValueError(f"positions must end in a dimension of 3, got {tuple(pos.shape)}")
    pos = pos.reshape(-1, 3)
    if f64:
        forward_only(mesh, pos, weight)
        w, ws = _weight(weight, pos.shape[0], dtype=dt)
        return _ops.paint_f64(mesh.clone(), pos, w, ws)
    w, ws = _weight(weight, pos.shape[0])
    return _autograd.CicPaint.apply(mesh, pos, w, ws, False, deterministic)


def cic_read(grid_mesh, positions, halo_size=0, sharding=None, method="direct"):
    """Absolute CIC read (painting.py:109-128 / 78-106); returns
    ``positions.shape[:-1]``.  A stack of meshes (nc, nx, ny, nz) returns
    ``positions.shape[:-1] + (nc,)``.  Additive keyword ``method``:
    "direct" (8 corner loads per particle) or "sorted" (a tile sort, then LDS
    windows; faster for particles in no spatial order, same values)."""
    single_device(sharding)
    f64 = wants_f64(grid_mesh, positions)
    dt = torch.float64 if f64 else torch.float32
    mesh = to_device(grid_mesh, dt)
    pos = to_device(positions, dt)
    if pos.shape[-1] != 3:
        raise ValueError(f"positions must end in a dimension of 3, got {tuple(pos.shape)}")
    out_shape = pos.shape[:-1]
    multi = mesh.dim() == 4
    m4 = mesh if multi else mesh.unsqueeze(0)
    if f64:
        forward_only(m4, pos)
        vals = _ops.read_f64(m4.contiguous(), pos.reshape(-1, 3))
    else:
        vals = _autograd.CicRead.apply(m4, pos.reshape(-1, 3), False, False, method)
    if multi:
        return vals.reshape(*out_shape, m4.shape[0])
    return vals.reshape(out_shape)


def cic_paint_dx(displacements, halo_size=0, sharding=None, weight=1.0, chunk_size=2 ** 24,
                 deterministic=False):
    """Relative CIC paint (painting.py:192-215 / 161-189): particle (i,j,k)
    sits at its lattice node plus ``displacements[i,j,k]``; returns a fresh
    mesh of the lattice shape.  ``chunk_size`` is accepted for signature
    compatibility; the kernel needs no chunking."""
    single_device(sharding)
    f64 = wants_f64(displacements, weight)
    dt = torch.float64 if f64 else torch.float32
    disp = to_device(displacements, dt)
    if disp.dim() != 4 or disp.shape[-1] != 3:
        raise ValueError(f"displacements must be (nx, ny, nz, 3), got {tuple(disp.shape)}")
    shape = tuple(disp.shape[:-1])
    if is_scalar(weight):
        w, ws = None, scalar_value(weight)
    else:
        w = to_device(weight, dt)
        if tuple(w.shape) != shape:
            raise ValueError("Weight shape must match particle shape")
        w, ws = w.reshape(-1), 1.0
    if f64:
        forward_only(disp, w)
        mesh = torch.empty(shape, dtype=torch.float64, device=disp.device)
        return _ops.paint_f64(mesh, disp.reshape(-1, 3), w, ws, relative=True, zero=True)
    zeros = torch.zeros(shape, dtype=torch.float32, device=disp.device)
    return _autograd.CicPaint.apply(zeros, disp.reshape(-1, 3), w, ws, True, deterministic)


def cic_read_dx(grid_mesh, disp, halo_size=0, sharding=None):
    """Relative CIC read (painting.py:239-260 / 218-236): values at lattice
    node + ``disp``; returns the mesh shape."""
    single_device(sharding)
    f64 = wants_f64(grid_mesh, disp)
    dt = torch.float64 if f64 else torch.float32
    mesh = to_device(grid_mesh, dt)
    d = to_device(disp, dt)
    shape = tuple(mesh.shape[-3:])
    if d.numel() != 3 * shape[0] * shape[1] * shape[2]:
        raise ValueError(f"disp {tuple(d.shape)} does not match mesh {shape}")
    multi = mesh.dim() == 4
    m4 = mesh if multi else mesh.unsqueeze(0)
    if f64:
        forward_only(m4, d)
        vals = _ops.read_f64(m4.contiguous(), d.reshape(-1, 3), relative=True)
    else:
        vals = _autograd.CicRead.apply(m4, d.reshape(-1, 3), True, False)
    if multi:
        return vals.reshape(*shape, m4.shape[0])
    return vals.reshape(shape)


def cic_paint_2d(mesh, positions, weight):
    """Weighted 2-D CIC paint (painting.py:131-158): ``mesh`` [nx, ny] plus
    the deposits of ``positions`` (..., 2), each times ``weight`` (one value
    per particle, or None)."""
    m = to_device(mesh)
    if m.dim() != 2:
        raise ValueError(f"mesh must be 2-D, got shape {tuple(m.shape)}")
    pos = to_device(positions)
    if pos.shape[-1] != 2:
        raise ValueError(f"positions must end in a dimension of 2, got {tuple(pos.shape)}")
    pos = pos.reshape(-1, 2)
    w = None
    if weight is not None:
        w = to_device(weight).reshape(-1)
        if w.numel() != pos.shape[0]:
            raise ValueError("Weight shape must match particle shape")
    return _autograd.CicPaint2d.apply(m, pos, w)


def compensate_cic(field):
    """painting.py:263-275: divide by the CIC window, Re IFFT[FFT(field)
    prod_d sinc(k_d/2pi)^-2]."""
    f = to_device(field)
    return _autograd.Compensate.apply(f)
