"""Pins of the oracle's 2-D paint, Gaussian smoothing and density plane
(painting.py:131-158, utils.py:208-222, lensing.py:11-44) by properties
and library routines, not by restating their formulas."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import lensing as ol
from oracle import painting as op
from oracle import utils as ou


def _pos2(r, n, shape):
    # positions across several periods, negative ones included
    return r.uniform(-1.5, 2.5, size=(n, 2)) * np.array(shape)


def test_paint_2d_transpose_of_bilinear_interpolation():
    """<paint(x, w), g> = sum_p w_p g(x_p), g(x) the periodic bilinear
    interpolation of scipy.ndimage.map_coordinates (order 1, grid-wrap):
    pins the corner offsets, weights and wrap against an independent
    routine."""
    r = np.random.default_rng(0)
    shape = (7, 5)
    pos = _pos2(r, 200, shape)
    w = r.normal(size=200)
    g = r.normal(size=shape)
    lhs = np.sum(op.cic_paint_2d(np.zeros(shape), pos, w) * g)
    interp = ndimage.map_coordinates(g, pos.T, order=1, mode="grid-wrap")
    assert abs(lhs - np.sum(w * interp)) < 1e-10 * np.abs(w * interp).sum()


def test_paint_2d_mass_lattice_shift_and_accumulation():
    r = np.random.default_rng(1)
    shape = (6, 8)
    pos = _pos2(r, 100, shape)
    w = r.uniform(0.5, 2.0, 100)
    m = op.cic_paint_2d(np.zeros(shape), pos, w)
    assert np.isclose(m.sum(), w.sum())
    # a particle on a node deposits everything there, wrapped
    m1 = op.cic_paint_2d(np.zeros(shape), np.array([[-1.0, 9.0]]), None)
    want = np.zeros(shape)
    want[5, 1] = 1.0
    assert np.array_equal(m1, want)
    # an integer shift rolls the mesh
    ms = op.cic_paint_2d(np.zeros(shape), pos + np.array([2.0, -3.0]), w)
    assert np.allclose(ms, np.roll(m, (2, -3), axis=(0, 1)))
    # the given mesh accumulates (scatter_add into it)
    base = r.normal(size=shape)
    assert np.allclose(op.cic_paint_2d(base, pos, w), base + m)


def test_gaussian_smoothing_plane_waves_and_mean():
    """A cosine of wave numbers (a, b) / n is scaled by
    exp(-2 pi^2 sigma^2 (a^2 + b^2) / n^2); the mean is kept."""
    n, sigma = 32, 1.7
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for a, b in [(1, 0), (0, 3), (2, 5), (16, 16)]:
        im = np.cos(2 * np.pi * (a * i + b * j) / n)
        got = ou.gaussian_smoothing(im, sigma)
        assert np.allclose(got, im * np.exp(-2 * np.pi ** 2 * sigma ** 2 * (a * a + b * b) / n ** 2), atol=1e-12)
    im = np.random.default_rng(2).normal(size=(n, n)) + 3.0
    assert np.isclose(ou.gaussian_smoothing(im, sigma).mean(), im.mean())
    with pytest.raises(ValueError):
        ou.gaussian_smoothing(np.zeros((8, 6)), 1.0)


def test_density_plane_selection_wrap_and_normalisation():
    nx, res, width, center = 16, 8, 4.0, 6.0
    r = np.random.default_rng(3)
    pos = np.concatenate([r.uniform(0, nx, (500, 2)), r.uniform(0, nx, (500, 1))], 1)
    plane = ol.density_plane(pos, (nx, nx, nx), center, width, res)
    inside = (pos[:, 2] > center - width / 2) & (pos[:, 2] <= center + width / 2)
    norm = (nx / res) ** 2 * width
    assert np.isclose(plane.sum() * norm, inside.sum())
    # the slab is (lo, hi]: a particle at lo is out, one at hi is in
    edge = np.array([[1.0, 1.0, center - width / 2], [3.0, 3.0, center + width / 2]])
    pe = ol.density_plane(edge, (nx, nx, nx), center, width, res)
    assert np.isclose(pe.sum() * norm, 1.0) and pe[1, 1] > 0 and pe[0, 0] == 0
    # (x, y) wrap by nx: shifted copies paint the same plane
    shifted = pos + np.array([nx, -2 * nx, 0.0])
    assert np.allclose(ol.density_plane(shifted, (nx, nx, nx), center, width, res), plane)
    # a lattice of the plane's cells, all in the slab: a uniform plane
    g = (np.stack(np.meshgrid(np.arange(res), np.arange(res), indexing="ij"), -1).reshape(-1, 2) + 0.25) * nx / res
    lat = np.concatenate([g, np.full((res * res, 1), center)], 1)
    assert np.allclose(ol.density_plane(lat, (nx, nx, nx), center, width, res), 1.0 / norm)
