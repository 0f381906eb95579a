"""Binned power spectra on the GPU (jaxpm ``utils.py``).

``power_spectrum`` keeps the reference's signature and result:
``(k_avg, P)``, with ``k_avg`` a float64 NumPy array (host-side in the
reference too) and ``P`` a float32 device tensor, of shape (nbins,) for a
scalar ``multipoles`` and (len(multipoles), nbins) otherwise.  The field is
transformed once by hipFFT (R2C) and every mode, with its Hermitian mirror, is
binned by ``jpm_power_spectrum_f32``; the host only builds the bin edges and
normalises a few hundred sums.
"""
import numpy as np
import torch

from . import _ops

__all__ = ["power_spectrum", "transfer", "coherence", "pktranscoh"]


def _kedges(mesh_shape, box_shape, kedges):
    """Bin edges of utils.py:38-50 (``_initialize_pk``)."""
    kmax = np.pi * np.min(mesh_shape / box_shape)                      # utils.py:38
    if kedges is None or isinstance(kedges, (int, float)):
        if kedges is None:
            dk = 2 * np.pi / np.min(box_shape) * 2                     # utils.py:42-44
        elif isinstance(kedges, int):
            dk = kmax / (kedges + 1)                                   # utils.py:45-46
        else:
            dk = kedges                                                # utils.py:47-48
        kedges = np.arange(dk, kmax, dk) + dk / 2                      # utils.py:50
    kedges = np.asarray(kedges, np.float64).reshape(-1)
    if kedges.size > 1 and np.any(np.diff(kedges) < 0):
        raise ValueError("kedges must be increasing")
    return kedges


def _as_mesh(m, name):
    if not isinstance(m, torch.Tensor) or m.device.type != "cuda":
        raise TypeError(f"{name} must be a device tensor")
    if m.dim() != 3:
        raise ValueError(f"{name} must be a 3D mesh, got shape {tuple(m.shape)}")
    return m.to(torch.float32).contiguous()


def power_spectrum(mesh, mesh2=None, box_shape=None, kedges=None, multipoles=0, los=(0., 0., 1.)):
    """Auto (or cross, with ``mesh2``) power spectrum of 3D fields, with
    Legendre multipoles about the line of sight ``los`` — utils.py:76-128.

    Bins are np.digitize bins of the float32 |k| over ``kedges`` (None: width
    twice the fundamental; int: that many bins up to Nyquist; float: that
    width; or explicit edges), the first and last dropped.  ``P`` is
    normalised to (box/mesh)^3 per mode as the reference (utils.py:123);
    cross spectra are the modulus of the complex bin sums (utils.py:117-119).
    """
    mesh = _as_mesh(mesh, "mesh")
    mesh_shape = np.array(mesh.shape)
    box_shape = mesh_shape if box_shape is None else np.asarray(box_shape)
    if np.ndim(multipoles) == 0 and multipoles == 0:
        los = (0., 0., 0.)
    else:
        los = np.asarray(los, np.float64)
        los = los / np.linalg.norm(los)
    poles = [int(p) for p in np.atleast_1d(multipoles)]
    edges = _kedges(mesh_shape, box_shape, kedges)
    fac = [2 * np.pi * m / l for m, l in zip(mesh_shape, box_shape)]   # utils.py:52-53
    spec1 = _ops.r2c(mesh)
    spec2 = None
    if mesh2 is not None:
        mesh2 = _as_mesh(mesh2, "mesh2")
        if mesh2.shape != mesh.shape:
            raise ValueError("mesh and mesh2 differ in shape")
        spec2 = _ops.r2c(mesh2)
    sums = _ops.power_spectrum_sums(spec1, spec2, mesh.shape, fac, edges, poles, los).cpu().numpy()
    count, ksum = sums[0], sums[1]
    npoles = len(poles)
    m = float(np.prod(mesh_shape))
    re = sums[2:2 + npoles] / m                 # norm='ortho': |X|^2 / M
    if spec2 is None:
        psum = re
    else:
        im = sums[2 + npoles:2 + 2 * npoles] / m
        psum = (re ** 2 + im ** 2) ** .5
    with np.errstate(divide='ignore', invalid='ignore'):
        kavg = (ksum / count)[1:-1]
        pk = (psum / count)[:, 1:-1] * (box_shape / mesh_shape).prod()
    pk = torch.as_tensor(pk.astype(np.float32), device=mesh.device)
    return (kavg, pk[0]) if np.ndim(multipoles) == 0 else (kavg, pk)


def transfer(mesh0, mesh1, box_shape, kedges=None):
    """sqrt(P1 / P0) — utils.py:131-135."""
    ks, pk0 = power_spectrum(mesh0, box_shape=box_shape, kedges=kedges)
    ks, pk1 = power_spectrum(mesh1, box_shape=box_shape, kedges=kedges)
    return ks, (pk1 / pk0) ** .5


def coherence(mesh0, mesh1, box_shape, kedges=None):
    """P01 / sqrt(P0 P1) — utils.py:138-143."""
    ks, pk01 = power_spectrum(mesh0, mesh1, box_shape=box_shape, kedges=kedges)
    ks, pk0 = power_spectrum(mesh0, box_shape=box_shape, kedges=kedges)
    ks, pk1 = power_spectrum(mesh1, box_shape=box_shape, kedges=kedges)
    return ks, pk01 / (pk0 * pk1) ** .5


def pktranscoh(mesh0, mesh1, box_shape, kedges=None):
    """(k, P0, P1, transfer, coherence) — utils.py:146-151."""
    ks, pk01 = power_spectrum(mesh0, mesh1, box_shape=box_shape, kedges=kedges)
    ks, pk0 = power_spectrum(mesh0, box_shape=box_shape, kedges=kedges)
    ks, pk1 = power_spectrum(mesh1, box_shape=box_shape, kedges=kedges)
    return ks, pk0, pk1, (pk1 / pk0) ** .5, pk01 / (pk0 * pk1) ** .5


def gaussian_smoothing(im, sigma):
    """utils.py:208-222: the image filtered by exp(-|k|^2 / (2 s^2)), s =
    1 / (2 pi sigma), |k| the norm of the fftfreq wave numbers, through the
    library's 2-D transforms.  The reference builds its filter with
    ``meshgrid``'s 'xy' order, which broadcasts against the image only when
    it is square; non-square images raise ValueError here."""
    from . import _autograd
    from ._convert import to_device
    t = to_device(im)
    if t.dim() != 2 or t.shape[0] != t.shape[1]:
        raise ValueError(f"gaussian_smoothing takes a square 2-D image, got shape {tuple(t.shape)}")
    return _autograd.GaussianSmoothing.apply(t, float(sigma))
