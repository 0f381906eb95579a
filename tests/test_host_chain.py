"""Host-side chain rule of the differentiable simulation (CPU only)."""
import numpy as np
import torch

from jaxpm_amd import cosmology as pc
from jaxpm_amd import integrate, pm
from jaxpm_amd.diffsim import CosmoCoefficients


def test_central_difference_jacobian():
    def fn(x):
        return np.array([x[0] ** 2, x[0] * x[1], np.sin(x[1]), 3.0])

    p = torch.tensor([0.3, 0.8], dtype=torch.float64, requires_grad=True)
    y = CosmoCoefficients.apply(p, fn, 1e-4)
    g = torch.tensor([1.0, -2.0, 0.5, 7.0], dtype=torch.float64)
    (y * g).sum().backward()
    x0, x1 = 0.3, 0.8
    ref = [2 * x0 - 2.0 * x1, -2.0 * x0 + 0.5 * np.cos(x1)]
    np.testing.assert_allclose(p.grad.numpy(), ref, rtol=1e-7)


def test_cosmology_quantities_are_smooth():
    """The host quantities the chain rule differentiates (growth, P(k) table,
    leapfrog coefficients) have central differences that agree across step
    sizes: no solver noise."""
    base = pc.Cosmology(Omega_c=0.25, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)
    a = np.linspace(0.1, 1.0, 6)
    k = np.logspace(-4, 1, 32)

    def q(oc):
        c = base.replace(Omega_c=oc)
        k0, beta, gamma = integrate.kdk_schedule(c, a)
        return np.concatenate([pc.linear_matter_power(c, k), pm.lpt_coefficients(c, 0.1, 2), [k0], beta, gamma])

    d = [(q(0.25 + h) - q(0.25 - h)) / (2 * h) for h in (1e-4, 2e-5)]
    np.testing.assert_allclose(d[0], d[1], rtol=2e-6, atol=1e-12 * np.abs(d[0]).max())
    # P(k) is quadratic in sigma8 (linear_matter_power normalisation)
    c2 = base.replace(sigma8=1.6)
    np.testing.assert_allclose(pc.linear_matter_power(c2, k), 4 * pc.linear_matter_power(base, k), rtol=1e-12)


def test_autotune_ties_go_to_the_widest_chunk(tmp_path, monkeypatch):
    """Candidates within 3% of the fastest are tied; the widest chunk wins;
    among its slices, those within 1% of its fastest are tied and the
    narrowest wins; a clear winner keeps its place."""
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    from jaxpm_amd import autotune
    t = {(3, 2, 0, 0): 1.00, (5, 2, 0, 0): 1.025, (5, 3, 0, 1): 1.01, (4, 2, 0, 0): 0.99, (2, 2, 0, 0): 1.2}
    assert autotune._finish("tie", t)[0] == (5, 3, 0, 1)
    t = {(5, 1, 0, 0): 1.008, (5, 2, 0, 0): 1.00, (5, 3, 0, 0): 1.001, (4, 1, 0, 0): 0.99}
    assert autotune._finish("slice", t)[0] == (5, 1, 0, 0)
    assert autotune._finish("clear", {(3, 2, 0, 0): 1.0, (5, 2, 0, 0): 1.2})[0] == (3, 2, 0, 0)
