"""GPU lens planes (2-D paint, density plane, Gaussian smoothing) and their
adjoints against the oracle."""
import numpy as np
import pytest
import torch

from oracle import lensing as ol
from oracle import painting as op
from oracle import utils as ou

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.as_tensor(np.asarray(x, np.float32), device="cuda")


def rel_err(got, ref):
    return np.abs(np.asarray(got, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30)


@pytest.mark.parametrize("shape,weighted", [((64, 64), True), ((37, 50), False), ((512, 512), True)])
def test_cic_paint_2d_matches_oracle(gpu, shape, weighted):
    from jaxpm_amd import painting
    print(f"cic_paint_2d {shape} weighted={weighted}", flush=True)
    r = np.random.default_rng(10)
    n = 20000
    pos = (r.uniform(-1.2, 2.2, size=(n, 2)) * np.array(shape)).astype(np.float32)
    w = r.uniform(0.2, 2.0, n).astype(np.float32) if weighted else None
    base = r.normal(size=shape).astype(np.float32)
    got = painting.cic_paint_2d(dev(base), dev(pos), None if w is None else dev(w)).cpu().numpy()
    ref = op.cic_paint_2d(base, pos, w)
    assert rel_err(got, ref) < 2e-6
    with pytest.raises(ValueError):
        painting.cic_paint_2d(dev(base), dev(pos), dev(np.ones(n - 1)))


def test_cic_paint_2d_gradients(gpu):
    """Mesh, weight and position cotangents: the weight cotangent is the
    periodic bilinear read of the plane cotangent; the position cotangent
    matches float64 central differences of the oracle along random
    directions."""
    from scipy import ndimage

    from jaxpm_amd import painting
    r = np.random.default_rng(11)
    shape, n = (24, 20), 300
    pos = (r.uniform(-0.5, 1.5, size=(n, 2)) * np.array(shape)).astype(np.float32)
    w = r.uniform(0.5, 1.5, n).astype(np.float32)
    g = r.normal(size=shape)
    mt, pt, wt = dev(np.zeros(shape)).requires_grad_(True), dev(pos).requires_grad_(True), dev(w).requires_grad_(True)
    (painting.cic_paint_2d(mt, pt, wt) * dev(g)).sum().backward()
    assert np.allclose(mt.grad.cpu().numpy(), g.astype(np.float32))
    read = ndimage.map_coordinates(g, pos.astype(np.float64).T, order=1, mode="grid-wrap")
    assert rel_err(wt.grad.cpu().numpy(), read) < 1e-5
    gp = pt.grad.cpu().numpy().astype(np.float64)
    for k in range(4):
        u = r.normal(size=pos.shape) * 1e-4
        f = lambda x: np.sum(op.cic_paint_2d(np.zeros(shape), x, w.astype(np.float64)) * g)  # noqa: E731
        fd = (f(pos + u) - f(pos - u)) / 2
        assert abs(np.sum(gp * u) - fd) < 2e-3 * np.abs(gp * u).sum()


@pytest.mark.parametrize("res,sigma", [(64, None), (128, 1.5), (100, 2.0)])
def test_density_plane_matches_oracle(gpu, res, sigma):
    from jaxpm_amd import lensing
    print(f"density_plane res={res} sigma={sigma}", flush=True)
    r = np.random.default_rng(12)
    nx = 64
    pos = r.uniform(-0.3, 1.3, size=(60000, 3)).astype(np.float32) * nx
    got = lensing.density_plane(dev(pos), (nx, nx, nx), 30.0, 10.0, res, smoothing_sigma=sigma).cpu().numpy()
    ref = ol.density_plane(pos, (nx, nx, nx), 30.0, 10.0, res, smoothing_sigma=sigma, dtype=np.float32)
    assert got.shape == (res, res)
    assert rel_err(got, ref.astype(np.float64)) < 3e-6


@pytest.mark.parametrize("n", [64, 100, 512])
def test_gaussian_smoothing_matches_oracle(gpu, n):
    from jaxpm_amd import utils
    im = np.random.default_rng(13).normal(size=(n, n)).astype(np.float32)
    got = utils.gaussian_smoothing(dev(im), 1.3).cpu().numpy()
    assert rel_err(got, ou.gaussian_smoothing(im, 1.3)) < 3e-6
    with pytest.raises(ValueError):
        utils.gaussian_smoothing(dev(np.zeros((8, 16))), 1.0)


def test_density_plane_gradient(gpu):
    """Position cotangent of the smoothed density plane against float64
    central differences of the oracle along random directions (particles
    kept away from the slab faces, where the selection jumps)."""
    from jaxpm_amd import lensing
    r = np.random.default_rng(14)
    nx, res, center, width = 32, 32, 16.0, 8.0
    pos = r.uniform(0, nx, size=(3000, 3)).astype(np.float32)
    pos[:, 2] = np.where(r.uniform(size=3000) < 0.5, r.uniform(12.5, 19.5, 3000), r.uniform(21, 30, 3000))
    g = r.normal(size=(res, res))
    pt = dev(pos).requires_grad_(True)
    (lensing.density_plane(pt, (nx, nx, nx), center, width, res, smoothing_sigma=1.0) * dev(g)).sum().backward()
    gp = pt.grad.cpu().numpy().astype(np.float64)
    assert np.all(gp[:, 2] == 0)
    f = lambda x: np.sum(ol.density_plane(x, (nx, nx, nx), center, width, res, smoothing_sigma=1.0) * g)  # noqa: E731
    for k in range(4):
        u = r.normal(size=pos.shape) * 1e-4
        fd = (f(pos + u) - f(pos - u)) / 2
        assert abs(np.sum(gp * u) - fd) < 2e-3 * np.abs(gp * u).sum()
