"""Float64 paths under the x64 switch (the reference's equivalence tests run
with jax_enable_x64, tests/test_distributed_pm.py:30): cic_paint, cic_read,
cic_paint_dx, cic_read_dx and pm_forces in float64 against the float64
oracle at double-precision tolerances, and the float32 default for float64
input when the switch is off (as in JAX)."""
import warnings

import numpy as np
import pytest
import torch

from oracle import painting as op
from oracle import pm as opm

pytestmark = pytest.mark.gpu


@pytest.fixture
def x64():
    from jaxpm_amd import config
    config.update("jax_enable_x64", True)
    yield
    config.update("jax_enable_x64", False)


def lattice(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1)


def rel(got, ref):
    return np.abs(np.asarray(got) - ref).max() / np.abs(ref).max()


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 20, 9)])
def test_paint_read_f64(gpu, x64, shape):
    from jaxpm_amd import painting
    r = np.random.default_rng(1)
    n = 5000
    pos = r.uniform(-4, max(shape) + 4, (n, 3))
    w = r.normal(size=n)
    mesh0 = r.normal(size=shape)
    got = painting.cic_paint(torch.as_tensor(mesh0, device=gpu), torch.as_tensor(pos, device=gpu),
                             torch.as_tensor(w, device=gpu))
    assert got.dtype == torch.float64
    assert rel(got.cpu().numpy(), op.cic_paint(mesh0, pos, w)) < 1e-13
    m = r.normal(size=shape)
    rd = painting.cic_read(torch.as_tensor(m, device=gpu), torch.as_tensor(pos, device=gpu))
    assert rd.dtype == torch.float64
    assert rel(rd.cpu().numpy(), op.cic_read(m, pos)) < 1e-14


def test_relative_f64(gpu, x64):
    from jaxpm_amd import painting
    shape = (16, 8, 12)
    r = np.random.default_rng(2)
    disp = r.normal(scale=1.3, size=(*shape, 3))
    w = r.uniform(0.5, 1.5, shape)
    got = painting.cic_paint_dx(torch.as_tensor(disp, device=gpu), weight=torch.as_tensor(w, device=gpu))
    assert got.dtype == torch.float64
    assert rel(got.cpu().numpy(), op.cic_paint_dx(disp, w)) < 1e-13
    m = r.normal(size=shape)
    rd = painting.cic_read_dx(torch.as_tensor(m, device=gpu), torch.as_tensor(disp, device=gpu))
    assert rel(rd.cpu().numpy(), op.cic_read_dx(m, disp)) < 1e-14


@pytest.mark.parametrize("shape,absolute", [((16, 16, 16), True), ((12, 16, 20), True), ((16, 16, 16), False)])
def test_pm_forces_f64(gpu, x64, shape, absolute):
    from jaxpm_amd import pm
    r = np.random.default_rng(3)
    d = r.normal(scale=0.7, size=(*shape, 3))
    pos = lattice(shape) + d if absolute else d
    got = pm.pm_forces(torch.as_tensor(pos, device=gpu), mesh_shape=shape, paint_absolute_pos=absolute)
    assert got.dtype == torch.float64
    ref = opm.pm_forces(pos, mesh_shape=shape, paint_absolute_pos=absolute)
    assert rel(got.cpu().numpy(), ref) < 1e-12


def test_pm_forces_f64_options(gpu, x64):
    """Order-0 gradient, long-range split and a given real density."""
    from jaxpm_amd import pm
    shape = (16, 16, 16)
    r = np.random.default_rng(4)
    pos = lattice(shape) + r.normal(scale=0.7, size=(*shape, 3))
    got = pm.pm_forces(torch.as_tensor(pos, device=gpu), mesh_shape=shape, r_split=1.5, grad_order=0)
    pots = opm.force_potentials(np.fft.fftn(op.cic_paint(np.zeros(shape), pos)), 1.5, order=0)
    ref = np.stack([op.cic_read(p, pos) for p in pots], -1)
    assert rel(got.cpu().numpy(), ref) < 1e-12
    rho = r.normal(size=shape)
    got = pm.pm_forces(torch.as_tensor(pos, device=gpu), delta=torch.as_tensor(rho, device=gpu))
    ref = opm.pm_forces(pos, delta=rho)
    assert rel(got.cpu().numpy(), ref) < 1e-12


def test_f64_is_forward_only(gpu, x64):
    from jaxpm_amd import painting
    pos = torch.zeros((10, 3), dtype=torch.float64, device=gpu, requires_grad=True)
    with pytest.raises(NotImplementedError):
        painting.cic_paint(torch.zeros((8, 8, 8), dtype=torch.float64, device=gpu), pos)


def test_float64_input_without_x64_runs_float32(gpu):
    """With the switch off, float64 input is computed in float32 (JAX's
    default), with a warning the first time."""
    from jaxpm_amd import _convert, config, painting
    assert not config.x64_enabled()
    _convert._warned_f64 = False
    pos = np.random.default_rng(5).uniform(0, 8, (100, 3))
    with warnings.catch_warnings(record=True) as wl:
        warnings.simplefilter("always")
        out = painting.cic_paint(torch.zeros((8, 8, 8), device=gpu), pos)
    assert out.dtype == torch.float32
    assert any("jax_enable_x64" in str(w.message) for w in wl)
