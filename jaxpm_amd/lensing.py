"""Lens planes, mirroring ``jaxpm/lensing.py`` (``density_plane``).

The density plane of a slab of the box is one fused HIP pass over the
particles (``jpm_density_plane_f32``): the periodic wrap and rescaling of
(x, y), the slab selection, the 2-D CIC paint and the normalisation; the
optional smoothing runs the library's 2-D transforms.  Both are
differentiable with hand-written adjoints.  ``convergence_Born`` (jax_cosmo
distances, ``map_coordinates``) is not part of this build.
"""
from . import _autograd
from ._convert import to_device
from .utils import gaussian_smoothing


def density_plane(positions, box_shape, center, width, plane_resolution, smoothing_sigma=None):
    """lensing.py:11-44: the particles with center - width/2 < z <= center +
    width/2 painted onto a ``plane_resolution``^2 plane at (x, y) =
    mod(xy, nx) / nx * plane_resolution (both axes wrapped and scaled by the
    box's nx, as there), divided by (nx/res)(ny/res) width, then smoothed by
    ``smoothing_sigma`` pixels when given."""
    pos = to_device(positions)
    if pos.shape[-1] != 3:
        raise ValueError(f"positions must end in a dimension of 3, got {tuple(pos.shape)}")
    nx, ny, _ = (int(s) for s in box_shape)
    res = int(plane_resolution)
    plane = _autograd.DensityPlane.apply(pos.reshape(-1, 3), (nx, ny), float(center), float(width), res)
    if smoothing_sigma is not None:
        plane = gaussian_smoothing(plane, smoothing_sigma)
    return plane
