"""The two largest benchmark workloads run whole inside the GPU suite:

* 48 batched simulations, 16 each at 64^3, 96^3 and 128^3 (one particle per
  cell, box 4n Mpc/h, Omega_c ~ U[0.2, 0.35] and sigma8 ~ U[0.7, 0.9] from
  default_rng(6), 1LPT at a = 0.1 + 20 leapfrog steps, checkpoints every 5,
  the time loop replayed from captured HIP graphs): mass of every final
  density, finite gradients, the first simulation of each class against the
  per-simulation path, and the first 64^3 simulation against the oracle's
  float64 forward model (loss and central differences along noise
  directions).
* 512^3 particles on a 1024^3 mesh (box 1000 Mpc/h, 2LPT + 30 steps,
  checkpoints every 5): mass conservation and a finite gradient at full
  size, and the same two-grid arrangement (particles on a lattice of
  spacing 2 cells) scaled down to 32^3 on 64^3 against the oracle's forces,
  force VJP and leapfrog adjoint.

The reference runs many simulations with jax.vmap (tests/
test_distributed_pm.py:335-409); it has no 1024^3 test.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint as oa
from oracle import cosmology as oc
from oracle import ode as oo
from oracle import pm as opm

pytestmark = pytest.mark.gpu

KTAB = np.logspace(-4, 1, 128)
SIZES = [64] * 16 + [96] * 16 + [128] * 16
STEPS = np.linspace(0.1, 1.0, 21)


def _cosmo_params():
    r = np.random.default_rng(6)
    om = r.uniform(0.2, 0.35, len(SIZES))
    s8 = r.uniform(0.7, 0.9, len(SIZES))
    return [dict(Omega_c=float(a), Omega_b=0.05, h=0.7, n_s=0.96, sigma8=float(b)) for a, b in zip(om, s8)]


def _noise(b, n):
    return np.random.default_rng(1000 + b).standard_normal((n, n, n), dtype=np.float32)


def _target(b, n):
    return np.random.default_rng(2000 + b).normal(scale=0.3, size=(n, n, n)).astype(np.float32)


@pytest.fixture(scope="module")
def batch48(gpu):
    """One forward + adjoint pass of the 48-simulation batch; returns the
    final densities and noise gradients of every simulation (host arrays of
    the first simulation of each class, plus per-simulation summaries)."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import diffsim
    print("config 4: 48 batched simulations", flush=True)
    torch.cuda.empty_cache()
    params = _cosmo_params()
    shapes = [(n, n, n) for n in SIZES]
    boxes = [(4.0 * n,) * 3 for n in SIZES]
    model = diffsim.BatchLptNBody(shapes, boxes, STEPS, [pc.Cosmology(**p) for p in params], order=1,
                                  checkpoint_every=5, graph=True)
    noise = torch.cat([torch.as_tensor(_noise(b, n), device=gpu).reshape(-1) for b, n in enumerate(SIZES)])
    target = torch.cat([torch.as_tensor(_target(b, n), device=gpu).reshape(-1) for b, n in enumerate(SIZES)])
    nt = noise.requires_grad_(True)
    rho = model.forward_fixed_cosmology(nt)
    (((rho - 1.0 - target) ** 2).sum() / 0.1 ** 2).backward()
    torch.cuda.synchronize()
    offs = np.concatenate([[0], np.cumsum([n ** 3 for n in SIZES])])
    out = {"mass": [], "finite": [], "grad_rms": [], "rho": {}, "grad": {}}
    for b, n in enumerate(SIZES):
        r = rho[offs[b]:offs[b + 1]]
        g = nt.grad[offs[b]:offs[b + 1]]
        out["mass"].append(float(r.double().sum()))
        out["finite"].append(bool(torch.isfinite(g).all()) and bool(torch.isfinite(r).all()))
        out["grad_rms"].append(float(g.double().square().mean().sqrt()))
        if b % 16 == 0:
            out["rho"][b] = r.detach().reshape(n, n, n).cpu().numpy()
            out["grad"][b] = g.reshape(n, n, n).cpu().numpy()
    out["params"] = params
    del model, rho, nt, noise, target
    torch.cuda.empty_cache()
    return out


def test_batch48_mass_and_finite_gradients(batch48):
    for b, n in enumerate(SIZES):
        assert abs(batch48["mass"][b] - n ** 3) < 1e-6 * n ** 3, (b, batch48["mass"][b])
        assert batch48["finite"][b], b
        assert batch48["grad_rms"][b] > 0, b


@pytest.mark.parametrize("b", [0, 16, 32])
def test_batch48_matches_per_simulation(gpu, batch48, b):
    """The first simulation of each class through the per-simulation path
    (the same model, one simulation, no graphs)."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import diffsim
    n = SIZES[b]
    shape = (n, n, n)
    single = diffsim.LptNBody(shape, (4.0 * n,) * 3, STEPS, base_cosmo=pc.Cosmology(**batch48["params"][b]),
                              order=1, checkpoint_every=5)
    nt = torch.as_tensor(_noise(b, n), device=gpu).requires_grad_(True)
    rho = single(nt, single.parameter_vector())
    (((rho - 1.0 - torch.as_tensor(_target(b, n), device=gpu)) ** 2).sum() / 0.1 ** 2).backward()
    rb, gb = batch48["rho"][b], batch48["grad"][b]
    r1, g1 = rho.detach().cpu().numpy(), nt.grad.cpu().numpy()
    rho_rms = np.sqrt(np.mean((rb - r1) ** 2)) / np.sqrt(np.mean((r1 - 1) ** 2))
    g_rms = np.sqrt(np.mean((gb - g1) ** 2)) / np.sqrt(np.mean(g1 ** 2))
    print(f"class {n}^3: density rms rel diff {rho_rms:.2e}, gradient rms rel diff {g_rms:.2e}", flush=True)
    assert rho_rms < 1e-4
    assert g_rms < 1e-3


def _oracle_density(noise, n, params):
    c = oc.Cosmology(**params)
    shape, box = (n, n, n), (4.0 * n,) * 3
    pk = oc.linear_matter_power(c, KTAB)
    ic = opm.linear_field(shape, box, lambda k: np.interp(k, KTAB, pk), noise)
    parts = opm.uniform_particles(shape).astype(np.float64)
    dx, p, _ = opm.lpt(c, ic, parts, a=STEPS[0], order=1)
    x, _ = oo.leapfrog((parts + dx).reshape(-1, 3), p.reshape(-1, 3), STEPS, c, shape, reuse_force=True)
    return oo.final_density(x, shape)


def test_batch48_64_against_oracle(batch48):
    """Simulation 0 (64^3) against the oracle's float64 forward model: the
    final density, the loss and the noise gradient along two random
    directions (float64 central differences)."""
    b, n = 0, 64
    params = batch48["params"][b]
    noise = _noise(b, n).astype(np.float64)
    target = _target(b, n).astype(np.float64)
    print("config 4: oracle forward model at 64^3", flush=True)
    ref = _oracle_density(noise, n, params)
    rho = batch48["rho"][b]
    d_rms = np.sqrt(np.mean((rho - ref) ** 2)) / np.sqrt(np.mean((ref - 1) ** 2))
    loss = np.sum((rho.astype(np.float64) - 1 - target) ** 2) / 0.1 ** 2
    ref_loss = np.sum((ref - 1 - target) ** 2) / 0.1 ** 2
    print(f"density rms rel diff {d_rms:.2e}; loss {loss:.6e} vs oracle {ref_loss:.6e}", flush=True)
    assert d_rms < 1e-3
    assert abs(loss - ref_loss) < 1e-4 * ref_loss
    g = batch48["grad"][b].astype(np.float64)
    r = np.random.default_rng(8)
    eps = 1e-6
    for i in range(2):
        d = r.normal(size=noise.shape)
        lp = np.sum((_oracle_density(noise + eps * d, n, params) - 1 - target) ** 2) / 0.1 ** 2
        lm = np.sum((_oracle_density(noise - eps * d, n, params) - 1 - target) ** 2) / 0.1 ** 2
        fd = (lp - lm) / (2 * eps)
        an = float(np.sum(g * d))
        print(f"direction {i}: adjoint {an:.6e}, central difference {fd:.6e}", flush=True)
        assert abs(an - fd) < 1e-3 * abs(fd), (an, fd)


def test_large_1024_mesh_512_particles(gpu):
    """The full-size two-grid run: one forward + adjoint with mass
    conservation of the final density and a finite, nonzero gradient."""
    from jaxpm_amd import _autograd, diffsim, integrate
    from jaxpm_amd import cosmology as pc
    print("config 5: 512^3 particles on 1024^3, 2LPT + 30 steps", flush=True)
    torch.cuda.empty_cache()
    nm, npart, S = 1024, 512, 30
    shape = (nm, nm, nm)
    spec = diffsim.SimSpec(shape, (1000.0,) * 3, np.linspace(0.1, 1.0, S + 1),
                           base_cosmo=pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8), order=2)
    gen = torch.Generator(device=gpu).manual_seed(5)
    nt = torch.randn(shape, generator=gen, device=gpu).requires_grad_(True)
    ax = torch.arange(npart, device=gpu, dtype=torch.float32) * 2.0
    lattice = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1)
    ktab = torch.as_tensor(spec.ktab, dtype=torch.float32, device=gpu)
    stepper = integrate.Stepper(shape, npart ** 3)
    q, x0, v0 = spec.initial_state(nt, spec.parameter_vector(), True, lattice, ktab)
    del lattice
    xS, _ = integrate.NBody.apply(x0, v0, q[spec.slices["kdk"]], stepper, 5)
    rho = _autograd.CicPaint.apply(torch.zeros(shape, device=gpu), xS, None, 1.0, False, False)
    mass = float(rho.detach().double().sum())
    mean = npart ** 3 / nm ** 3
    loss = ((rho / mean - 1.0) ** 2).sum() / 0.1 ** 2
    loss.backward()
    torch.cuda.synchronize()
    g = nt.grad
    finite = bool(torch.isfinite(g).all())
    rms = float(g.double().square().mean().sqrt())
    peak = torch.cuda.max_memory_allocated() / 1e9
    print(f"mass {mass:.6e} (want {npart ** 3}), loss {float(loss):.6e}, grad rms {rms:.3e}, finite {finite}, "
          f"peak torch {peak:.1f} GB", flush=True)
    del stepper, xS, rho, loss, q, x0, v0, nt, g
    torch.cuda.empty_cache()
    assert abs(mass - npart ** 3) < 1e-6 * npart ** 3
    assert finite and rms > 0


def _two_grid_state(seed):
    """32^3 particles on a lattice of spacing 2 cells of a 64^3 mesh, with
    displacements and velocities."""
    r = np.random.default_rng(seed)
    ax = np.arange(32) * 2.0
    lat = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    x = (lat + r.normal(scale=1.2, size=lat.shape)).astype(np.float32)
    v = r.normal(scale=0.3, size=lat.shape).astype(np.float32)
    return x, v


def test_two_grid_forces_and_vjp_against_oracle(gpu):
    from jaxpm_amd import integrate
    shape = (64, 64, 64)
    x, _ = _two_grid_state(40)
    u = np.random.default_rng(41).normal(size=x.shape).astype(np.float32)
    st = integrate.Stepper(shape, x.shape[0])
    xt = torch.as_tensor(x, device=gpu)
    f = torch.empty_like(xt)
    field = st.new_field(keep_sort=True)
    st.forces(xt, f, field)
    ref = opm.pm_forces(x.astype(np.float64), mesh_shape=shape)
    err = np.abs(f.cpu().numpy() - ref).max() / np.abs(ref).max()
    g = torch.zeros_like(xt)
    st.vjp_field(field, xt, torch.as_tensor(u, device=gpu), 0.7, g)
    gref = 0.7 * oa.force_vjp(x, u, shape)
    gerr = np.abs(g.cpu().numpy() - gref).max() / np.abs(gref).max()
    print(f"two-grid 32^3 on 64^3: forces {err:.2e}, force VJP {gerr:.2e}", flush=True)
    assert err < 2e-5
    assert gerr < 5e-5


def test_two_grid_leapfrog_adjoint_against_oracle(gpu):
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    shape = (64, 64, 64)
    x0, v0 = _two_grid_state(42)
    steps = np.linspace(0.3, 0.7, 6)
    r = np.random.default_rng(43)
    gx, gv = r.normal(size=x0.shape).astype(np.float32), r.normal(size=x0.shape).astype(np.float32)
    xt = torch.as_tensor(x0, device=gpu).requires_grad_(True)
    vt = torch.as_tensor(v0, device=gpu).requires_grad_(True)
    x, v = integrate.leapfrog(xt, vt, steps, pc.Cosmology(), shape, checkpoint_every=2)
    ((x * torch.as_tensor(gx, device=gpu)).sum() + (v * torch.as_tensor(gv, device=gpu)).sum()).backward()
    rx, rv = oo.leapfrog(x0, v0, steps, oc.Cosmology(), shape)
    lx, lv = oa.leapfrog_vjp(x0, v0, steps, oc.Cosmology(), shape, gx, gv)
    ex = np.abs(x.detach().cpu().numpy() - rx).max() / np.abs(rx - x0).max()
    elx = np.abs(xt.grad.cpu().numpy() - lx).max() / np.abs(lx).max()
    elv = np.abs(vt.grad.cpu().numpy() - lv).max() / np.abs(lv).max()
    print(f"two-grid leapfrog: positions {ex:.2e}, adjoint x {elx:.2e}, v {elv:.2e}", flush=True)
    assert ex < 1e-4
    assert elx < 2e-4 and elv < 2e-4
