"""Fourier-space multipliers, mirroring ``jaxpm/kernels.py``.

These public helpers return the multiplier arrays for callers that compose
their own pipelines (elementwise torch expressions over broadcast wave-number
vectors).  The PM step itself never materialises them: the fused k-space
kernels of libjaxpm_hip.so evaluate the same formulas from per-axis tables.
"""
import math

import torch

from . import _lib


def _shape_of(k_array):
    if isinstance(k_array, (tuple, list)) and all(isinstance(s, int) for s in k_array):
        return tuple(k_array)
    return tuple(k_array.shape)


def fftk(k_array):
    """kernels.py:10-23: [kx (nx,1,1), ky (1,ny,1), kz (1,1,nz)] with
    k = 2*pi*fftfreq(n) in radians per cell.  ``k_array`` is the (c2c)
    spectrum or a mesh shape."""
    shape = _shape_of(k_array)
    if len(shape) != 3:
        raise ValueError(f"fftk needs a 3-D array, got shape {shape}")
    dev = _lib.device()
    out = []
    for d, n in enumerate(shape):
        v = (2 * math.pi * torch.fft.fftfreq(n, dtype=torch.float64)).to(torch.float32).to(dev)
        s = [1, 1, 1]
        s[d] = n
        out.append(v.reshape(s))
    return out


def gradient_kernel(kvec, direction, order=1):
    """kernels.py:41-66: ``i k`` (order 0, entry len//2 zeroed) or the
    4-point difference ``i (8 sin w - sin 2w) / 6`` (order 1)."""
    w = kvec[direction]
    if order == 0:
        flat = w.reshape(-1).clone()
        flat[len(flat) // 2] = 0
        return (1j * flat.to(torch.complex64)).reshape(w.shape)
    a = 1 / 6.0 * (8 * torch.sin(w) - torch.sin(2 * w))
    return a.to(torch.complex64) * 1j


def invlaplace_kernel(kvec, fd=False):
    """kernels.py:69-92: -1/kk with 0 at k = 0."""
    if fd:
        kk = sum((ki * torch.sinc(ki / (2 * math.pi))) ** 2 for ki in kvec)
    else:
        kk = sum(ki ** 2 for ki in kvec)
    kk_nozeros = torch.where(kk == 0, torch.ones_like(kk), kk)
    return -torch.where(kk == 0, torch.zeros_like(kk), 1 / kk_nozeros)


def longrange_kernel(kvec, r_split):
    """kernels.py:95-115: exp(-kk r^2), or 1 when r_split == 0."""
    if r_split != 0:
        kk = sum(ki ** 2 for ki in kvec)
        return torch.exp(-kk * r_split ** 2)
    return 1.


def cic_compensation(kvec):
    """kernels.py:118-136: prod_d sinc(k_d / 2pi)^-2 (normalised sinc)."""
    kwts = [torch.sinc(kvec[i] / (2 * math.pi)) for i in range(3)]
    return (kwts[0] * kwts[1] * kwts[2]) ** (-2)
