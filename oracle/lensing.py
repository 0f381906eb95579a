"""Oracle lens-plane density (test infrastructure only).

Restates JaxPM ``jaxpm/lensing.py:11-44`` (``density_plane``) with NumPy on
the CPU, on top of :func:`oracle.painting.cic_paint_2d` and
:func:`oracle.utils.gaussian_smoothing`.
"""
import numpy as np

from .painting import cic_paint_2d
from .utils import gaussian_smoothing


def density_plane(positions, box_shape, center, width, plane_resolution, smoothing_sigma=None, dtype=np.float64):
    """lensing.py:11-44: the particles with center - width/2 < z <=
    center + width/2, their (x, y) wrapped by ``jnp.mod(xy, nx)`` and scaled
    by ``plane_resolution / nx`` (both axes by nx, lensing.py:26-29), painted
    onto a zero plane, divided by (nx/res)(ny/res) width, then optionally
    smoothed."""
    nx, ny, nz = box_shape
    pos = np.asarray(positions, dtype).reshape(-1, 3)
    xy = np.mod(pos[:, :2], dtype(nx))
    xy = xy / dtype(nx) * dtype(plane_resolution)
    d = pos[:, 2]
    if dtype == np.float32:
        lo, hi = np.float32(center - width / 2), np.float32(center + width / 2)
    else:
        lo, hi = center - width / 2, center + width / 2
    weight = np.where((d > lo) & (d <= hi), dtype(1), dtype(0))
    plane = cic_paint_2d(np.zeros((plane_resolution, plane_resolution), dtype), xy, weight, dtype)
    norm = (nx / plane_resolution) * (ny / plane_resolution) * width
    plane = plane / (np.float32(norm) if dtype == np.float32 else norm)
    if smoothing_sigma is not None:
        plane = gaussian_smoothing(plane, smoothing_sigma).astype(dtype)
    return plane
