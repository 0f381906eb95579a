"""The kernel microbenchmark workload at full size, against the float64 oracle:
256^3 particles on a 256^3 mesh (box 1024 Mpc/h), in the two position states
it names -- uniform random (default_rng(1)) and clustered (seed-2 white noise,
Eisenstein-Hu P(k) with Omega_c = 0.25, Omega_b = 0.05, h = 0.7, n_s = 0.96,
sigma8 = 0.8, 2LPT at a = 0.1, then 20 leapfrog steps to a = 1, particles in
Lagrangian order) -- cic_paint into zeros, cic_read of the painted density,
and pm_forces on the clustered state.  The reference's own tests compare
these functions against FastPM/pmesh at 32^3-64^3 (tests/test_against_fpm.py);
the oracle stands in for the reference here."""
import numpy as np
import pytest
import torch

from oracle import painting as op
from oracle import pm as opm

pytestmark = pytest.mark.gpu

N1 = 256
SHAPE = (N1, N1, N1)


def _clustered(gpu):
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate, pm
    cosmo = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    k, pk = pc.pk_table(cosmo)
    gen = torch.Generator(device=gpu).manual_seed(2)
    with torch.no_grad():
        ic = pm.linear_field(SHAPE, (1024.0,) * 3, pm.TabulatedPk(k, pk),
                             torch.randn(SHAPE, generator=gen, device=gpu))
        lattice = torch.stack(torch.meshgrid(*[torch.arange(N1, device=gpu, dtype=torch.float32)] * 3,
                                             indexing="ij"), -1)
        dx, p, _ = pm.lpt(cosmo, ic, lattice, a=0.1, order=2)
        x0 = (lattice + dx).reshape(-1, 3).contiguous()
        v0 = p.reshape(-1, 3).contiguous()
        x, _ = integrate.leapfrog(x0, v0, np.linspace(0.1, 1.0, 21), cosmo, SHAPE)
    return x.reshape(-1, 3).contiguous()


@pytest.fixture(scope="module")
def states(gpu):
    print("config 2: uniform and clustered 256^3 states", flush=True)
    uni = torch.as_tensor(np.random.default_rng(1).uniform(0, N1, (N1 ** 3, 3)).astype(np.float32), device=gpu)
    clu = _clustered(gpu)
    torch.cuda.synchronize()
    return {"uniform": uni, "clustered": clu}


@pytest.mark.parametrize("state", ["uniform", "clustered"])
def test_paint_and_read_256(gpu, states, state):
    from jaxpm_amd import painting
    pos = states[state]
    print(f"config 2 paint/read, {state}", flush=True)
    rho = painting.cic_paint(torch.zeros(SHAPE, device=gpu), pos)
    got = rho.double().cpu().numpy()
    p64 = pos.double().cpu().numpy()
    ref = op.cic_paint(np.zeros(SHAPE), p64)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    rms = np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean((ref - ref.mean()) ** 2))
    print(f"paint: max error {err:.3e} of max, rms {rms:.3e} of the density contrast's", flush=True)
    assert err < 2e-6 and rms < 1e-6
    assert abs(got.sum() - N1 ** 3) < 1e-6 * N1 ** 3
    # the read of the painted density at the same positions (the oracle reads
    # the GPU's density, so only the read's own rounding is compared)
    rd_t = painting.cic_read(rho, pos)
    rd = rd_t.double().cpu().numpy()
    refr = op.cic_read(got, p64)
    err_r = np.abs(rd - refr).max() / np.abs(refr).max()
    print(f"read: max error {err_r:.3e} of max", flush=True)
    assert err_r < 2e-6
    # the tile-sorted read gives the same bits as the direct one
    assert torch.equal(painting.cic_read(rho, pos, method="sorted"), rd_t)


def test_pm_forces_256_clustered(gpu, states):
    from jaxpm_amd import pm
    pos = states["clustered"]
    print("config 2 pm_forces, clustered", flush=True)
    f = pm.pm_forces(pos, mesh_shape=SHAPE).double().cpu().numpy()
    ref = opm.pm_forces(pos.double().cpu().numpy(), mesh_shape=SHAPE)
    err = np.abs(f - ref).max() / np.abs(ref).max()
    rms = np.sqrt(np.mean((f - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    print(f"pm_forces: max error {err:.3e} of max, rms {rms:.3e}", flush=True)
    assert err < 2e-5 and rms < 2e-6
