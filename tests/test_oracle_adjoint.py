"""Pins for oracle.adjoint by float64 central differences."""
import numpy as np

from oracle import adjoint as oa
from oracle import cosmology as oc
from oracle import ode as oo
from oracle import painting as op
from oracle import pm as opm


def lattice(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)


def test_read_grad_fd():
    r = np.random.default_rng(0)
    mesh = r.normal(size=(6, 5, 4))
    pos = r.uniform(-3, 9, size=(40, 3))
    pos = np.floor(pos) + np.clip(pos - np.floor(pos), 0.01, 0.99)
    g = oa.cic_read_grad(mesh, pos)
    eps = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = eps
        fd = (op.cic_read(mesh, pos + e) - op.cic_read(mesh, pos - e)) / (2 * eps)
        np.testing.assert_allclose(g[:, d], fd, atol=1e-7)


def test_force_vjp_fd():
    shape = (8, 8, 8)
    r = np.random.default_rng(1)
    pos = lattice(shape) + r.normal(scale=0.8, size=(512, 3))
    u = r.normal(size=(512, 3))
    g = oa.force_vjp(pos, u, shape)
    eps = 1e-6
    for _ in range(4):
        d = r.normal(size=pos.shape)
        fd = (np.sum(opm.pm_forces(pos + eps * d, mesh_shape=shape) * u) -
              np.sum(opm.pm_forces(pos - eps * d, mesh_shape=shape) * u)) / (2 * eps)
        assert abs(np.sum(g * d) - fd) < 1e-5 * (abs(fd) + 1e-6)


def test_leapfrog_vjp_fd():
    shape = (8, 8, 8)
    c = oc.Cosmology()
    r = np.random.default_rng(2)
    x0 = lattice(shape) + r.normal(scale=0.5, size=(512, 3))
    v0 = r.normal(scale=0.2, size=(512, 3))
    steps = np.linspace(0.2, 0.5, 4)
    gx, gv = r.normal(size=(512, 3)), r.normal(size=(512, 3))
    lx, lv = oa.leapfrog_vjp(x0, v0, steps, c, shape, gx, gv)

    def loss(x, v):
        xs, vs = oo.leapfrog(x, v, steps, c, shape)
        return np.sum(gx * xs) + np.sum(gv * vs)

    eps = 1e-6
    for _ in range(3):
        dx, dv = r.normal(size=x0.shape), r.normal(size=v0.shape)
        fd = (loss(x0 + eps * dx, v0 + eps * dv) - loss(x0 - eps * dx, v0 - eps * dv)) / (2 * eps)
        an = np.sum(lx * dx) + np.sum(lv * dv)
        assert abs(an - fd) < 1e-5 * (abs(fd) + 1e-6), (an, fd)


def test_cpu_timing_port_matches_oracle():
    """The float32 SciPy port timed as the CPU baseline computes the same step."""
    from oracle import cpu_step
    shape = (16, 16, 16)
    r = np.random.default_rng(3)
    n = 16 ** 3
    x = (lattice(shape) + r.normal(scale=0.7, size=(n, 3))).astype(np.float32)
    v = r.normal(scale=0.1, size=(n, 3)).astype(np.float32)
    cx, cv = r.normal(size=(n, 3)).astype(np.float32), r.normal(size=(n, 3)).astype(np.float32)
    x1, v1, lx, lv = cpu_step.step_fwd_adj(x, v, cx, cv, shape, 0.01, 0.02, 2)
    xr = x + 0.01 * v
    np.testing.assert_allclose(x1, xr, atol=1e-6)
    vr = v + 0.02 * opm.pm_forces(xr, mesh_shape=shape)
    assert np.abs(v1 - vr).max() < 1e-6 * np.abs(vr).max() + 1e-7
    lxr = cx + oa.force_vjp(xr, 0.02 * cv.astype(np.float64), shape)
    assert np.abs(lx - lxr).max() < 1e-6 * np.abs(lxr).max()
    np.testing.assert_allclose(lv, cv + 0.01 * lx, atol=1e-6)
