"""Oracle binned power spectra (test infrastructure only).

Restates JaxPM ``jaxpm/utils.py:14-150``: ``_initialize_pk``,
``power_spectrum`` (auto and cross spectra, Legendre multipoles), and the
``transfer`` / ``coherence`` / ``pktranscoh`` compositions.

The binning follows the reference's rounding: ``|k|`` is the float64 sum of
squares cast to float32 and square-rooted there (``jnp.sqrt`` of a float64
NumPy array with x64 off, utils.py:54), then ``np.digitize`` against the
float64 edges (utils.py:56).  Modes of integer-valued ``|k|`` fall exactly
on the default edges (odd multiples of the fundamental), so this decides
which bin they land in.  Power sums are float64 here, float32 in the
reference.
"""
import numpy as np


def _legendre(ell, mu):
    """Legendre polynomial P_ell(mu) by Bonnet's recurrence."""
    p0, p1 = np.ones_like(mu), mu
    if ell == 0:
        return p0
    for n in range(1, ell):
        p0, p1 = p1, ((2 * n + 1) * mu * p1 - n * p0) / (n + 1)
    return p1


def initialize_pk(mesh_shape, box_shape, kedges, los):
    """utils.py:14-73.  Returns (dig, kcount, kavg, mumesh, kedges)."""
    mesh_shape = np.asarray(mesh_shape)
    box_shape = np.asarray(box_shape)
    kmax = np.pi * np.min(mesh_shape / box_shape)                      # utils.py:38
    if kedges is None or isinstance(kedges, (int, float)):
        if kedges is None:
            dk = 2 * np.pi / np.min(box_shape) * 2                     # utils.py:42-44
        elif isinstance(kedges, int):
            dk = kmax / (kedges + 1)                                   # utils.py:45-46
        else:
            dk = kedges                                                # utils.py:47-48
        kedges = np.arange(dk, kmax, dk) + dk / 2                      # utils.py:50
    kedges = np.asarray(kedges, np.float64)
    kvec = []
    for d, (m, l) in enumerate(zip(mesh_shape, box_shape)):
        s = [1, 1, 1]
        s[d] = int(m)
        kvec.append((2 * np.pi * m / l) * np.fft.fftfreq(int(m)).reshape(s))   # utils.py:52-53
    ksq = sum(ki ** 2 for ki in kvec)
    kmesh = np.sqrt(np.broadcast_to(ksq, tuple(mesh_shape)).astype(np.float32))   # utils.py:54
    dig = np.digitize(kmesh.reshape(-1), kedges)                       # utils.py:56
    kcount = np.bincount(dig, minlength=len(kedges) + 1)
    kavg = np.bincount(dig, weights=kmesh.reshape(-1), minlength=len(kedges) + 1) / kcount
    kavg = kavg[1:-1]
    if los is None:
        mumesh = 1.0
    else:
        mun = sum(ki * li for ki, li in zip(kvec, los))
        safe = np.where(kmesh == 0, 1, kmesh)
        mumesh = np.where(kmesh == 0, 0, mun / safe)                   # utils.py:67-71
    return dig, kcount, kavg, mumesh, kedges


def power_spectrum(mesh, mesh2=None, box_shape=None, kedges=None, multipoles=0, los=(0., 0., 1.)):
    """utils.py:76-128.  Returns (k_avg, P), P of shape (len(poles), nbins)
    or (nbins,) for a scalar ``multipoles``."""
    mesh = np.asarray(mesh, np.float64)
    mesh_shape = np.array(mesh.shape)
    box_shape = mesh_shape if box_shape is None else np.asarray(box_shape)
    if np.ndim(multipoles) == 0 and multipoles == 0:
        los = None
    else:
        los = np.asarray(los, np.float64)
        los = los / np.linalg.norm(los)
    poles = np.atleast_1d(multipoles)
    dig, kcount, kavg, mumesh, _ = initialize_pk(mesh_shape, box_shape, kedges, los)
    n_bins = len(kavg) + 2
    meshk = np.fft.fftn(mesh, norm='ortho')
    if mesh2 is None:
        mmk = meshk.real ** 2 + meshk.imag ** 2
    else:
        mmk = meshk * np.fft.fftn(np.asarray(mesh2, np.float64), norm='ortho').conj()
    pk = np.empty((len(poles), n_bins))
    for i, ell in enumerate(poles):
        w = (mmk * (2 * ell + 1) * _legendre(int(ell), np.broadcast_to(mumesh, mesh.shape))).reshape(-1)
        if mesh2 is None:
            psum = np.bincount(dig, weights=w, minlength=n_bins)[:n_bins]
        else:
            re = np.bincount(dig, weights=w.real, minlength=n_bins)[:n_bins]
            im = np.bincount(dig, weights=w.imag, minlength=n_bins)[:n_bins]
            psum = (re ** 2 + im ** 2) ** .5                           # utils.py:117-119
        pk[i] = psum
    with np.errstate(divide='ignore', invalid='ignore'):
        pk = (pk / kcount)[:, 1:-1] * (box_shape / mesh_shape).prod()  # utils.py:123
    return (kavg, pk[0]) if np.ndim(multipoles) == 0 else (kavg, pk)


def transfer(mesh0, mesh1, box_shape, kedges=None):
    """utils.py:131-135."""
    ks, pk0 = power_spectrum(mesh0, box_shape=box_shape, kedges=kedges)
    ks, pk1 = power_spectrum(mesh1, box_shape=box_shape, kedges=kedges)
    return ks, (pk1 / pk0) ** .5


def coherence(mesh0, mesh1, box_shape, kedges=None):
    """utils.py:138-143."""
    ks, pk01 = power_spectrum(mesh0, mesh1, box_shape=box_shape, kedges=kedges)
    ks, pk0 = power_spectrum(mesh0, box_shape=box_shape, kedges=kedges)
    ks, pk1 = power_spectrum(mesh1, box_shape=box_shape, kedges=kedges)
    return ks, pk01 / (pk0 * pk1) ** .5


def gaussian_smoothing(im, sigma):
    """utils.py:208-222: ifft2(fft2(im) * filter).real, filter =
    norm.pdf(|k|, 0, 1 / (2 pi sigma)) / its value at k = 0, |k| the norm of
    the fftfreq wave numbers (meshgrid in 'xy' order, utils.py:214-216,
    which broadcasts against the image only for square images)."""
    im = np.asarray(im, np.float64)
    if im.ndim != 2 or im.shape[0] != im.shape[1]:
        raise ValueError("gaussian_smoothing takes a square 2-D image")
    kx, ky = np.meshgrid(np.fft.fftfreq(im.shape[0]), np.fft.fftfreq(im.shape[1]))
    k = np.sqrt(kx ** 2 + ky ** 2)
    s = 1.0 / (2.0 * np.pi * sigma)
    filt = np.exp(-0.5 * (k / s) ** 2)
    return np.fft.ifft2(np.fft.fft2(im) * filt).real
