"""Oracle Fourier multipliers (test infrastructure only).

Restates JaxPM ``jaxpm/kernels.py:10-136`` in float64.  Wave numbers are in
radians per cell, ``2*pi*fftfreq(n)``, broadcastable per axis (``fftk``,
kernels.py:10-23, via ``jaxdecomp.fftfreq3d``).
"""
import numpy as np


def fftk(shape):
    """kernels.py:10-23: [kx (n0,1,1), ky (1,n1,1), kz (1,1,n2)]."""
    out = []
    for d, n in enumerate(shape):
        s = [1, 1, 1]
        s[d] = n
        out.append((2 * np.pi * np.fft.fftfreq(n)).reshape(s))
    return out


def gradient_kernel(kvec, direction, order=1):
    """kernels.py:41-66.

    order 0: ``i*k`` with entry ``len//2`` zeroed (57-61).
    order 1: ``i*(8 sin w - sin 2w)/6``, the 4-point difference (63-66).
    """
    w = kvec[direction]
    if order == 0:
        wts = np.array(1j * w)
        flat = wts.reshape(-1)
        flat[len(flat) // 2] = 0
        return flat.reshape(w.shape)
    return 1j * (8 * np.sin(w) - np.sin(2 * w)) / 6.0


def invlaplace_kernel(kvec, fd=False):
    """kernels.py:69-92: -1/kk with 0 at kk == 0; fd uses k*sinc(k/2pi)."""
    if fd:
        kk = sum((ki * np.sinc(ki / (2 * np.pi))) ** 2 for ki in kvec)
    else:
        kk = sum(ki ** 2 for ki in kvec)
    nz = np.where(kk == 0, 1, kk)
    return -np.where(kk == 0, 0, 1 / nz)


def longrange_kernel(kvec, r_split):
    """kernels.py:95-115: exp(-kk r^2), or the scalar 1 when r_split == 0."""
    if r_split != 0:
        kk = sum(ki ** 2 for ki in kvec)
        return np.exp(-kk * r_split ** 2)
    return 1.0


def cic_compensation(kvec):
    """kernels.py:118-136: prod_d sinc(k_d / 2pi)^-2 (normalised sinc)."""
    w = np.sinc(kvec[0] / (2 * np.pi)) * np.sinc(kvec[1] / (2 * np.pi)) * \
        np.sinc(kvec[2] / (2 * np.pi))
    return w ** (-2)
