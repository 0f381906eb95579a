"""NaN and +-Inf inputs propagate to the cells and results they touch, as the
reference's float scatter_add and gathers do (jaxpm/painting.py:39-44,
104-106), on every paint path: the binned fixed-point paint (single- and
multi-chunk tiles, absolute and relative), the direct deterministic paint,
the 3-channel paint of the force VJP and the leapfrog adjoint.  The finite
cells keep their values."""
import numpy as np
import pytest
import torch

from oracle import painting as op

pytestmark = pytest.mark.gpu


def lattice(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)


def _check_mesh(got, ref):
    """Same NaN cells, same +-Inf cells, finite cells to float rounding."""
    got = np.asarray(got, np.float64)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(np.isposinf(got), np.isposinf(ref))
    np.testing.assert_array_equal(np.isneginf(got), np.isneginf(ref))
    fin = np.isfinite(ref)
    assert np.abs(got[fin] - ref[fin]).max() <= 2e-6 * np.abs(ref[fin]).max()


def _weights(n, seed):
    r = np.random.default_rng(seed)
    w = r.uniform(0.5, 1.5, n).astype(np.float32)
    idx = r.choice(n, 5, replace=False)
    w[idx[0]] = np.nan
    w[idx[1]] = np.inf
    w[idx[2]] = -np.inf
    w[idx[3]] = np.inf     # two infinities of the same sign
    w[idx[4]] = np.inf
    return w


@pytest.mark.parametrize("shape,heavy", [((16, 16, 16), False), ((20, 24, 36), False), ((32, 32, 32), True)])
@pytest.mark.parametrize("method", ["binned", "atomic-deterministic"])
def test_paint_nonfinite_weights(gpu, shape, heavy, method):
    from jaxpm_amd import _ops
    r = np.random.default_rng(1)
    n = 20000
    pos = r.uniform(-3, max(shape) + 3, (n, 3)).astype(np.float32)
    if heavy:   # one tile split over several chunks (exact int64 windows)
        pos[: 3 * n // 4] = r.uniform(2, 13, (3 * n // 4, 3)).astype(np.float32)
    w = _weights(n, 2)
    print(f"nonfinite paint {method} {shape} heavy={heavy}", flush=True)
    mesh = torch.zeros(shape, device=gpu)
    if method == "binned":
        _ops.paint(mesh, torch.as_tensor(pos, device=gpu), torch.as_tensor(w, device=gpu), deterministic=True)
    else:
        _ops.paint(mesh, torch.as_tensor(pos, device=gpu), torch.as_tensor(w, device=gpu), deterministic=True,
                   method="atomic")
    ref = op.cic_paint(np.zeros(shape), pos.astype(np.float64), w.astype(np.float64))
    assert np.isnan(ref).any() and np.isinf(ref).any()
    _check_mesh(mesh.cpu().numpy(), ref)


def test_paint_nonfinite_scalar_weight(gpu):
    from jaxpm_amd import _ops
    shape = (16, 16, 16)
    pos = np.random.default_rng(3).uniform(0, 16, (1000, 3)).astype(np.float32)
    mesh = torch.zeros(shape, device=gpu)
    _ops.paint(mesh, torch.as_tensor(pos, device=gpu), None, wscalar=float("inf"), deterministic=True)
    ref = op.cic_paint(np.zeros(shape), pos.astype(np.float64), np.inf)
    _check_mesh(mesh.cpu().numpy(), ref)


def test_paint_dx_nonfinite_weights(gpu):
    """Relative painting (cic_paint_dx) with non-finite weights."""
    from jaxpm_amd import painting
    shape = (16, 16, 16)
    r = np.random.default_rng(4)
    disp = r.normal(scale=1.5, size=(*shape, 3)).astype(np.float32)
    w = _weights(int(np.prod(shape)), 5).reshape(shape)
    got = painting.cic_paint_dx(torch.as_tensor(disp, device=gpu), weight=torch.as_tensor(w, device=gpu))
    ref = op.cic_paint_dx(disp.astype(np.float64), w.astype(np.float64))
    _check_mesh(got.cpu().numpy(), ref)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_paint_nonfinite_position(gpu, bad):
    """A particle at a non-finite position deposits NaN (its corner weights
    are NaN) into some cells; every other cell keeps the value painted
    without that particle."""
    from jaxpm_amd import painting
    shape = (16, 16, 16)
    r = np.random.default_rng(6)
    pos = r.uniform(0, 16, (5000, 3)).astype(np.float32)
    pos[17, 1] = bad
    got = painting.cic_paint(torch.zeros(shape, device=gpu), torch.as_tensor(pos, device=gpu)).cpu().numpy()
    ref = op.cic_paint(np.zeros(shape), np.delete(pos, 17, 0).astype(np.float64))
    nan = np.isnan(got)
    assert 1 <= nan.sum() <= 8
    assert np.abs(got[~nan] - ref[~nan]).max() <= 2e-6 * ref.max()


def test_force_vjp_nan_cotangent_gives_nan_gradient(gpu):
    """A NaN in the cotangent of pm_forces reaches the whole position gradient
    (through the painted cotangent and the FFT), as jax.grad's does; before,
    the fixed-point paint turned it into a finite value."""
    from jaxpm_amd import pm
    shape = (16, 16, 16)
    r = np.random.default_rng(7)
    pos = (lattice(shape) + r.normal(scale=0.5, size=(4096, 3))).astype(np.float32)
    u = r.normal(size=pos.shape).astype(np.float32)
    u[100, 2] = np.nan
    pt = torch.as_tensor(pos, device=gpu).requires_grad_(True)
    f = pm.pm_forces(pt, mesh_shape=shape)
    (f * torch.as_tensor(u, device=gpu)).sum().backward()
    assert torch.isnan(pt.grad).all()


def test_leapfrog_adjoint_inf_cotangent(gpu):
    """An infinite cotangent of one final position reaches the initial
    gradients as non-finite values, through the checkpointed adjoint."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    shape = (16, 16, 16)
    r = np.random.default_rng(8)
    x0 = (lattice(shape) + r.normal(scale=0.5, size=(4096, 3))).astype(np.float32)
    v0 = r.normal(scale=0.2, size=x0.shape).astype(np.float32)
    gx = r.normal(size=x0.shape).astype(np.float32)
    gx[5, 0] = np.inf
    xt = torch.as_tensor(x0, device=gpu).requires_grad_(True)
    vt = torch.as_tensor(v0, device=gpu).requires_grad_(True)
    x, v = integrate.leapfrog(xt, vt, np.linspace(0.3, 0.6, 4), pc.Cosmology(), shape, checkpoint_every=2)
    (x * torch.as_tensor(gx, device=gpu)).sum().backward()
    assert not torch.isfinite(xt.grad).all()
    assert torch.isnan(vt.grad).any()


def test_finite_inputs_unchanged(gpu):
    """Finite inputs take the fixed-point path alone: the binned paint is
    still bitwise reproducible and equal to the direct deterministic one to
    float rounding."""
    from jaxpm_amd import _ops
    shape = (32, 32, 32)
    r = np.random.default_rng(9)
    pos = torch.as_tensor(r.uniform(0, 32, (50000, 3)).astype(np.float32), device=gpu)
    w = torch.as_tensor(r.normal(size=50000).astype(np.float32), device=gpu)
    a = _ops.paint(torch.zeros(shape, device=gpu), pos, w, deterministic=True)
    b = _ops.paint(torch.zeros(shape, device=gpu), pos, w, deterministic=True)
    c = _ops.paint(torch.zeros(shape, device=gpu), pos, w, deterministic=True, method="atomic")
    assert torch.equal(a, b)
    assert float((a - c).abs().max()) <= 1e-6 * float(a.abs().max())
