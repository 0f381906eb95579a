"""Single-GPU FFT facade, lattice and noise, mirroring ``jaxpm/distributed.py``.

The reference's sharding / halo machinery (distributed.py:17-34, 45-113)
reduces to the identity on one device; on one GPU periodic wrap happens
inside the CIC kernels, so only the single-device branches exist here.
"""
import torch

from . import _ops
from ._convert import single_device, to_device


def fft3d(x):
    """distributed.py:37-38: unnormalised c2c forward FFT (full spectrum).
    Real input goes through R2C and Hermitian extension."""
    if isinstance(x, torch.Tensor) and x.is_complex() or (not isinstance(x, torch.Tensor) and
                                                         hasattr(x, "dtype") and x.dtype.kind == "c"):
        return _ops.c2c(to_device(x, torch.complex64))
    xr = to_device(x)
    return _ops.half_to_full(_ops.r2c(xr), tuple(xr.shape))


def ifft3d(x):
    """distributed.py:41-42: ``pifft3d(x).real`` -- inverse c2c FFT with 1/M,
    real part, computed as C2R of the Hermitian part of ``x``."""
    xk = to_device(x, torch.complex64)
    shape = tuple(xk.shape)
    out = _ops.c2r(_ops.full_to_half(xk), shape, clobber=True)
    return out.mul_(1.0 / xk.numel())


def uniform_particles(mesh_shape, sharding=None):
    """distributed.py:168-190 (single device: 187-190): int32 lattice
    (nx, ny, nz, 3)."""
    single_device(sharding)
    from . import _lib
    dev = _lib.device()
    grids = torch.meshgrid(*[torch.arange(s, device=dev, dtype=torch.int32) for s in mesh_shape],
                           indexing="ij")
    return torch.stack(grids, dim=-1)


def normal_field(seed, shape, sharding=None, dtype=torch.float32):
    """distributed.py:193-223 (single device: 222-223): N(0,1) white noise.

    ``seed`` is an int or a ``torch.Generator`` on the device.  The stream is
    torch's Philox generator, not JAX's threefry: the same integer seed gives
    different noise than ``jax.random.normal``.  To reproduce a JAX run, pass
    the JAX noise array to :func:`jaxpm_amd.pm.linear_field` instead.
    """
    single_device(sharding)
    from . import _lib
    dev = _lib.device()
    if isinstance(seed, torch.Generator):
        gen = seed
    else:
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed))
    return torch.randn(tuple(shape), generator=gen, device=dev, dtype=dtype)
