"""Gradients of a density loss with respect to the white noise and to
(Omega_c, sigma8), through the hand-written adjoints, against float64
central differences of the oracle's forward model (linear_field -> lpt ->
leapfrog -> cic_paint, as the reference's tests/test_gradients.py:27-72)."""
import numpy as np
import pytest
import torch

from oracle import adjoint as oa
from oracle import cosmology as oc
from oracle import ode as oo
from oracle import pm as opm

pytestmark = pytest.mark.gpu

BASE = dict(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
KTAB = np.logspace(-4, 1, 128)


def dev(x):
    return torch.as_tensor(np.asarray(x, np.float32), device="cuda")


def oracle_density(noise, shape, box, a_steps, order, absolute, **cosmo):
    c = oc.Cosmology(**{**BASE, **cosmo})
    pk = oc.linear_matter_power(c, KTAB)
    ic = opm.linear_field(shape, box, lambda k: np.interp(k, KTAB, pk), noise)
    parts = opm.uniform_particles(shape).astype(np.float64) if absolute else None
    dx, p, _ = opm.lpt(c, ic, parts, a=a_steps[0], order=order)
    if absolute:
        x0, v0 = (parts + dx).reshape(-1, 3), p.reshape(-1, 3)
    else:
        x0, v0 = dx, p
    x, _ = oo.leapfrog(x0, v0, a_steps, c, shape, paint_absolute_pos=absolute)
    return oo.final_density(x, shape, paint_absolute_pos=absolute)


def loss_np(rho, target, sigma=0.1):
    return float(np.sum((rho - 1.0 - target) ** 2) / sigma ** 2)


@pytest.mark.parametrize("order,absolute", [(2, True), (1, True), (2, False)])
def test_gradients_noise_and_cosmology(gpu, order, absolute):
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import diffsim
    shape, box = (16, 16, 16), (64.,) * 3
    a_steps = np.linspace(0.1, 1.0, 7)
    print(f"diffsim order={order} absolute={absolute}", flush=True)
    noise = np.random.default_rng(3).normal(size=shape)
    target = oracle_density(np.random.default_rng(4).normal(size=shape), shape, box, a_steps, order, absolute) - 1.0
    model = diffsim.LptNBody(shape, box, a_steps, base_cosmo=pc.Cosmology(**BASE), order=order,
                             paint_absolute_pos=absolute, checkpoint_every=2)
    nt = dev(noise).requires_grad_(True)
    params = model.parameter_vector().requires_grad_(True)
    rho = model(nt, params)
    loss = ((rho - 1.0 - dev(target)) ** 2).sum() / 0.1 ** 2
    ref_loss = loss_np(oracle_density(noise, shape, box, a_steps, order, absolute), target)
    print(f"loss {float(loss.detach()):.8e} vs oracle {ref_loss:.8e}", flush=True)
    assert abs(float(loss.detach()) - ref_loss) < 1e-4 * ref_loss
    loss.backward()
    gn = nt.grad.cpu().numpy().astype(np.float64)
    gp = params.grad.numpy()

    def L(nz, **kw):
        return loss_np(oracle_density(nz, shape, box, a_steps, order, absolute, **kw), target)

    # steps small enough that almost no particle crosses a cell boundary
    # (where the CIC weights, hence the loss, have kinks) between the two
    # evaluations; float64 keeps the differences well above round-off
    r = np.random.default_rng(5)
    an, fd = [], []
    eps = 1e-6
    for _ in range(16 if (order, absolute) == (2, True) else 4):
        d = r.normal(size=shape)
        fd.append((L(noise + eps * d) - L(noise - eps * d)) / (2 * eps))
        an.append(float(np.sum(gn * d)))
    an, fd = np.array(an), np.array(fd)
    print(f"noise directions: max |adjoint - FD| / max |FD| = {np.abs(an - fd).max() / np.abs(fd).max():.2e}",
          flush=True)
    assert np.abs(an - fd).max() < 1e-3 * np.abs(fd).max(), (an, fd)
    for j, (nm, h) in enumerate((("Omega_c", 1e-6), ("sigma8", 1e-6))):
        v = BASE[nm]
        fdp = (L(noise, **{nm: v + h}) - L(noise, **{nm: v - h})) / (2 * h)
        print(f"{nm}: adjoint {gp[j]:.8e}, FD {fdp:.8e}, rel {abs(gp[j] - fdp) / abs(fdp):.2e}", flush=True)
        assert abs(gp[j] - fdp) < 1e-3 * abs(fdp), (nm, gp[j], fdp)


def test_leapfrog_coefficient_gradients(gpu):
    """d/d(k0, beta, gamma) of <gx, xS> + <gv, vS> against central
    differences of the oracle's leapfrog over explicit coefficients."""
    from jaxpm_amd import integrate
    shape = (16, 16, 16)
    r = np.random.default_rng(36)
    n = int(np.prod(shape))
    lat = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    x0 = (lat + r.normal(scale=0.6, size=(n, 3))).astype(np.float32)
    v0 = r.normal(scale=0.3, size=(n, 3)).astype(np.float32)
    gx, gv = r.normal(size=(n, 3)), r.normal(size=(n, 3))
    c = np.concatenate([[0.05], [0.3, 0.25, 0.2], [0.08, 0.06, 0.04]])
    st = integrate.Stepper(shape, n)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    xS, vS = integrate.NBody.apply(dev(x0), dev(v0), ct, st, 2)
    ((xS * dev(gx)).sum() + (vS * dev(gv)).sum()).backward()

    def L(cc):
        x, v = x0.astype(np.float64), v0.astype(np.float64)
        v = v + cc[0] * opm.pm_forces(x, mesh_shape=shape)
        for i in range(3):
            x = x + cc[1 + i] * v
            v = v + cc[4 + i] * opm.pm_forces(x, mesh_shape=shape)
        return float(np.sum(x * gx) + np.sum(v * gv))

    for j in range(len(c)):
        h = 1e-6
        e = np.zeros_like(c)
        e[j] = h
        fd = (L(c + e) - L(c - e)) / (2 * h)
        assert abs(ct.grad[j].item() - fd) < 1e-3 * (abs(fd) + 1e-2), (j, ct.grad[j].item(), fd)


def test_linear_field_table_gradient(gpu):
    from jaxpm_amd import _autograd
    shape, box = (16, 12, 20), (64., 48., 80.)
    r = np.random.default_rng(37)
    noise = r.normal(size=shape)
    g = r.normal(size=shape)
    pk = oc.linear_matter_power(oc.Cosmology(**BASE), KTAB)
    kt, pt = dev(KTAB), dev(pk).requires_grad_(True)
    out = _autograd.LinearField.apply(dev(noise), box, kt, pt)
    (out * dev(g)).sum().backward()
    got = pt.grad.cpu().numpy()

    def L(tab):
        return float(np.sum(g * opm.linear_field(shape, box, lambda k: np.interp(k, KTAB, tab), noise)))

    active = np.nonzero(np.abs(got) > 0)[0]
    assert len(active) > 10
    for j in active[:: max(1, len(active) // 12)]:
        h = 1e-4 * pk[j]
        e = np.zeros_like(pk)
        e[j] = h
        fd = (L(pk + e) - L(pk - e)) / (2 * h)
        assert abs(got[j] - fd) < 1e-3 * abs(fd) + 1e-6 * np.abs(got).max(), (j, got[j], fd)


def test_lincomb_and_dot(gpu):
    from jaxpm_amd import _autograd, _ops
    r = np.random.default_rng(38)
    a, b = r.normal(size=1000), r.normal(size=1000)
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    _ops.dot_into(out, 1, dev(a), dev(b))
    ref = float(np.sum(a.astype(np.float32).astype(np.float64) * b.astype(np.float32).astype(np.float64)))
    assert abs(out[1].item() - ref) < 1e-12 * np.sum(np.abs(a * b))
    assert out[0].item() == 0.0 and out[2].item() == 0.0
    c = torch.tensor([0.5, -2.0], dtype=torch.float64, requires_grad=True)
    xa, xb = dev(a).requires_grad_(True), dev(b).requires_grad_(True)
    y = _autograd.LinComb.apply(c, xa, xb)
    np.testing.assert_allclose(y.detach().cpu().numpy(), 0.5 * a - 2.0 * b, rtol=1e-5, atol=1e-6)
    w = r.normal(size=1000)
    (y * dev(w)).sum().backward()
    np.testing.assert_allclose(c.grad.numpy(), [np.sum(w * a), np.sum(w * b)], rtol=1e-5)
    np.testing.assert_allclose(xa.grad.cpu().numpy(), 0.5 * w, rtol=1e-6)
    np.testing.assert_allclose(xb.grad.cpu().numpy(), -2.0 * w, rtol=1e-6)


def test_coefficient_leapfrog_matches_cosmology_leapfrog(gpu):
    """The coefficient-gradient recompute (forces, then axpy) reproduces the
    fused force+kick trajectory: the adjoint of (x0, v0) agrees with and
    without the coefficient cotangent."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    shape = (16, 16, 16)
    r = np.random.default_rng(39)
    n = int(np.prod(shape))
    lat = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    x0 = (lat + r.normal(scale=0.6, size=(n, 3))).astype(np.float32)
    v0 = r.normal(scale=0.3, size=(n, 3)).astype(np.float32)
    gx = r.normal(size=(n, 3))
    sched = integrate.schedule_tensor(integrate.kdk_schedule(pc.Cosmology(), np.linspace(0.2, 0.6, 5)))
    grads = []
    for want in (False, True):
        st = integrate.Stepper(shape, n, deterministic=True)
        xt, vt = dev(x0).requires_grad_(True), dev(v0).requires_grad_(True)
        ct = sched.clone().requires_grad_(want)
        xS, _ = integrate.NBody.apply(xt, vt, ct, st, 2)
        (xS * dev(gx)).sum().backward()
        grads.append((xt.grad.cpu().numpy(), vt.grad.cpu().numpy()))
    np.testing.assert_array_equal(grads[0][0], grads[1][0])
    np.testing.assert_array_equal(grads[0][1], grads[1][1])
    lx, lv = oa.leapfrog_vjp(x0, v0, np.linspace(0.2, 0.6, 5), oc.Cosmology(), shape, gx, np.zeros_like(gx))
    assert np.abs(grads[0][0] - lx).max() < 2e-4 * np.abs(lx).max()
