"""GPU leapfrog and its checkpointed adjoint against the oracle."""
import numpy as np
import pytest
import torch

from oracle import adjoint as oa
from oracle import cosmology as oc
from oracle import ode as oo

pytestmark = pytest.mark.gpu


def lattice(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)


def rel_err(got, ref):
    return np.abs(got - ref).max() / np.abs(ref).max()


def _state(shape, seed):
    r = np.random.default_rng(seed)
    n = int(np.prod(shape))
    x0 = (lattice(shape) + r.normal(scale=0.6, size=(n, 3))).astype(np.float32)
    v0 = r.normal(scale=0.3, size=(n, 3)).astype(np.float32)
    return x0, v0


def test_force_vjp_matches_oracle(gpu):
    from jaxpm_amd import integrate
    shape = (16, 16, 16)
    x, _ = _state(shape, 30)
    u = np.random.default_rng(31).normal(size=x.shape).astype(np.float32)
    st = integrate.Stepper(shape, x.shape[0])
    xt = torch.as_tensor(x, device="cuda")
    phi = st.new_phi()
    st.field(xt, phi)
    g = torch.zeros_like(xt)
    st.vjp(phi, xt, torch.as_tensor(u, device="cuda"), 0.7, g)
    ref = 0.7 * oa.force_vjp(x, u, shape)
    assert rel_err(g.cpu().numpy(), ref) < 5e-5


@pytest.mark.parametrize("shape", [(16, 16, 16), (16, 16, 32)])
def test_leapfrog_forward(gpu, shape):
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    x0, v0 = _state(shape, 32)
    steps = np.linspace(0.1, 1.0, 6)
    x, v = integrate.leapfrog(torch.as_tensor(x0, device="cuda"), torch.as_tensor(v0, device="cuda"), steps,
                              pc.Cosmology(), shape)
    rx, rv = oo.leapfrog(x0, v0, steps, oc.Cosmology(), shape)
    assert rel_err(x.cpu().numpy() - x0, rx - x0) < 1e-4
    assert rel_err(v.cpu().numpy(), rv) < 1e-4


@pytest.mark.parametrize("checkpoint_every", [None, 1, 2, 3])
def test_leapfrog_adjoint(gpu, checkpoint_every):
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    shape = (16, 16, 16)
    x0, v0 = _state(shape, 33)
    steps = np.linspace(0.2, 0.6, 6)
    r = np.random.default_rng(34)
    gx, gv = r.normal(size=x0.shape).astype(np.float32), r.normal(size=x0.shape).astype(np.float32)
    xt = torch.as_tensor(x0, device="cuda").requires_grad_(True)
    vt = torch.as_tensor(v0, device="cuda").requires_grad_(True)
    x, v = integrate.leapfrog(xt, vt, steps, pc.Cosmology(), shape, checkpoint_every=checkpoint_every)
    ((x * torch.as_tensor(gx, device="cuda")).sum() + (v * torch.as_tensor(gv, device="cuda")).sum()).backward()
    lx, lv = oa.leapfrog_vjp(x0, v0, steps, oc.Cosmology(), shape, gx, gv)
    assert rel_err(xt.grad.cpu().numpy(), lx) < 2e-4
    assert rel_err(vt.grad.cpu().numpy(), lv) < 2e-4
