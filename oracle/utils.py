"""Oracle binned power spectrum (test infrastructure only).

Restates JaxPM ``jaxpm/utils.py:14-128`` (``_initialize_pk`` and
``power_spectrum``) for the auto spectrum, monopole only.
"""
import numpy as np


def power_spectrum(mesh, box_shape=None, kedges=None):
    """utils.py:76-128 with ``multipoles=0`` and ``mesh2=None``.

    Returns (k_avg, P) over bins [1:-1] of ``np.digitize(|k|, kedges)``.
    """
    mesh = np.asarray(mesh, np.float64)
    mesh_shape = np.array(mesh.shape)
    box_shape = mesh_shape.astype(np.float64) if box_shape is None else np.asarray(box_shape, np.float64)
    kmax = np.pi * np.min(mesh_shape / box_shape)                      # utils.py:38
    if kedges is None or isinstance(kedges, (int, float)):
        if kedges is None:
            dk = 2 * np.pi / np.min(box_shape) * 2                     # utils.py:42-44
        elif isinstance(kedges, int):
            dk = kmax / (kedges + 1)
        else:
            dk = kedges
        kedges = np.arange(dk, kmax, dk) + dk / 2                      # utils.py:50
    kvec = []
    for d, (m, l) in enumerate(zip(mesh_shape, box_shape)):
        s = [1, 1, 1]
        s[d] = m
        kvec.append((2 * np.pi * m / l) * np.fft.fftfreq(m).reshape(s))
    kmesh = np.sqrt(sum(ki ** 2 for ki in kvec))
    dig = np.digitize(kmesh.reshape(-1), kedges)
    kcount = np.bincount(dig, minlength=len(kedges) + 1)
    kavg = np.bincount(dig, weights=kmesh.reshape(-1), minlength=len(kedges) + 1) / kcount
    kavg = kavg[1:-1]
    n_bins = len(kavg) + 2
    meshk = np.fft.fftn(mesh, norm='ortho')
    mmk = meshk.real ** 2 + meshk.imag ** 2
    psum = np.bincount(dig, weights=mmk.reshape(-1), minlength=n_bins)[:n_bins]
    pk = (psum / kcount[:n_bins])[1:-1] * (box_shape / mesh_shape).prod()
    return kavg, pk
