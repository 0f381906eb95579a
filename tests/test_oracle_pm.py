"""Pins for oracle.kernels / oracle.pm / oracle.ode / oracle.utils by closed
forms, hand-computed values and invariants."""
import numpy as np
import pytest

from oracle import cosmology as oc
from oracle import kernels as ok
from oracle import ode as oo
from oracle import pm as opm
from oracle import utils as ou


def a_fd(w):
    return (8 * np.sin(w) - np.sin(2 * w)) / 6.0


def test_kernel_values_by_hand():
    kv = ok.fftk((4, 4, 4))
    np.testing.assert_allclose(kv[0].ravel(), [0, np.pi / 2, -np.pi, -np.pi / 2])
    g1 = ok.gradient_kernel(kv, 0, order=1).ravel()
    np.testing.assert_allclose(g1[1], 1j * 4 / 3)                 # (8*1 - 0)/6
    assert abs(g1[2]) < 1e-15                                       # sin(pi), sin(2pi)
    g0 = ok.gradient_kernel(kv, 0, order=0).ravel()
    np.testing.assert_allclose(g0, [0, 1j * np.pi / 2, 0, -1j * np.pi / 2])  # len//2 zeroed
    il = ok.invlaplace_kernel(kv)
    assert il[0, 0, 0] == 0
    np.testing.assert_allclose(il[1, 0, 0], -1 / (np.pi / 2) ** 2)
    np.testing.assert_allclose(il[1, 1, 0], -1 / (2 * (np.pi / 2) ** 2))
    ilfd = ok.invlaplace_kernel(kv, fd=True)
    # k sinc(k/2pi) = 2 sin(k/2): at k = pi/2 -> 2 sin(pi/4) = sqrt(2)
    np.testing.assert_allclose(ilfd[1, 0, 0], -1 / 2.0)
    assert ok.longrange_kernel(kv, 0) == 1.0
    np.testing.assert_allclose(ok.longrange_kernel(kv, 2.0)[2, 0, 0], np.exp(-np.pi ** 2 * 4))
    cc = ok.cic_compensation(kv)
    np.testing.assert_allclose(cc[2, 0, 0], (np.pi / 2) ** 2)       # sinc(1/2)^-2
    np.testing.assert_allclose(cc[0, 0, 0], 1.0)


@pytest.mark.parametrize("order", [0, 1])
def test_plane_wave_force(order):
    """delta = cos(k x) gives F_x = -(g(k)/k^2) sin(k x) at the nodes."""
    n, m = 16, 3
    k = 2 * np.pi * m / n
    x = np.arange(n)
    delta = np.cos(k * x)[:, None, None] * np.ones((n, n, n))
    pots = opm.force_potentials(np.fft.fftn(delta), order=order)
    g = k if order == 0 else a_fd(k)
    expect = -(g / k ** 2) * np.sin(k * x)[:, None, None] * np.ones((n, n, n))
    np.testing.assert_allclose(pots[0], expect, atol=1e-12)
    np.testing.assert_allclose(pots[1], 0, atol=1e-12)
    if order == 1:  # pm_forces uses the default 4-point gradient (pm.py:55)
        lattice = opm.uniform_particles((n, n, n)).reshape(-1, 3)
        f = opm.pm_forces(lattice, delta=delta)
        np.testing.assert_allclose(f[:, 0], expect.reshape(-1), atol=1e-12)


def test_longrange_damps_force():
    n, m = 16, 2
    k = 2 * np.pi * m / n
    delta = np.cos(k * np.arange(n))[:, None, None] * np.ones((n, n, n))
    p0 = opm.force_potentials(np.fft.fftn(delta))[0]
    p1 = opm.force_potentials(np.fft.fftn(delta), r_split=1.5)[0]
    np.testing.assert_allclose(p1, p0 * np.exp(-k ** 2 * 1.5 ** 2), atol=1e-12)


def test_second_order_source_two_modes():
    n = 16
    k1, k2 = 2 * np.pi * 2 / n, 2 * np.pi * 3 / n
    A, B = 0.7, -0.4
    x = np.arange(n)
    delta = A * np.cos(k1 * x)[:, None, None] + B * np.cos(k2 * x)[None, :, None] + np.zeros((n, n, n))
    d2 = opm.second_order_source(np.fft.fftn(delta))
    expect = (a_fd(k1) ** 2 / k1 ** 2) * (a_fd(k2) ** 2 / k2 ** 2) * A * B * \
        np.cos(k1 * x)[:, None, None] * np.cos(k2 * x)[None, :, None] * np.ones((n, n, n))
    np.testing.assert_allclose(d2, expect, atol=1e-12)


def eds():
    return oc.Cosmology(Omega_c=0.95, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=0.8)


def test_zeldovich_single_mode_eds():
    n, m, a = 16, 1, 0.2
    k = 2 * np.pi * m / n
    x = np.arange(n)
    ic = 0.1 * np.cos(k * x)[:, None, None] * np.ones((n, n, n))
    dx, p, f = opm.lpt(eds(), ic, a=a, order=1)
    psi = -(a_fd(k) / k ** 2) * 0.1 * np.sin(k * x)[:, None, None] * np.ones((n, n, n))
    # EdS: D1 = a, f1 = 1, E = a^-1.5
    np.testing.assert_allclose(dx[..., 0], a * psi, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(p[..., 0], a ** 2 * 1.0 * a ** -1.5 * a * psi, rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(dx[..., 1:], 0, atol=1e-12)


def test_linear_field_constant_pk_is_scaled_noise():
    shape, box = (8, 8, 16), (64., 64., 128.)
    w = np.random.default_rng(0).normal(size=shape)
    P0 = 3.0
    out = opm.linear_field(shape, box, lambda k: P0 * np.ones_like(k), w)
    np.testing.assert_allclose(out, w * np.sqrt(P0 * np.prod(shape) / np.prod(box)), atol=1e-12)


def test_kdk_free_drift():
    """A uniform lattice with a uniform velocity feels no force: positions
    move by v * sum h / (am^3 E(am))."""
    c = oc.Cosmology()
    shape = (8, 8, 8)
    x0 = opm.uniform_particles(shape).reshape(-1, 3).astype(np.float64)
    v0 = np.tile([0.3, -0.2, 0.1], (x0.shape[0], 1))
    steps = np.linspace(0.1, 1.0, 6)
    x, v = oo.leapfrog(x0, v0, steps, c, shape)
    drift = sum(h / (am ** 3 * np.sqrt(oc.Esqr(c, am)))
                for h, am in zip(np.diff(steps), steps[:-1] + 0.5 * np.diff(steps)))
    np.testing.assert_allclose(x, x0 + v0 * drift, atol=1e-9)
    np.testing.assert_allclose(v, v0, atol=1e-9)


def test_kdk_time_reversible():
    c = oc.Cosmology()
    shape = (8, 8, 8)
    r = np.random.default_rng(1)
    x0 = opm.uniform_particles(shape).reshape(-1, 3) + r.normal(scale=0.3, size=(512, 3))
    v0 = r.normal(scale=0.05, size=(512, 3))
    steps = np.linspace(0.3, 0.6, 4)
    x1, v1 = oo.leapfrog(x0, v0, steps, c, shape)
    x2, v2 = oo.leapfrog(x1, v1, steps[::-1], c, shape)
    np.testing.assert_allclose(x2, x0, atol=1e-9)
    np.testing.assert_allclose(v2, v0, atol=1e-9)


def brute_pk(mesh, box, mesh2=None, poles=(0,), los=(0., 0., 1.), dk=None):
    """Direct DFT by its definition, one mode at a time, binned by a scan
    over the edges; Legendre polynomials in closed form."""
    n = np.array(mesh.shape)
    L = np.array(box, np.float64)
    grid = np.stack(np.meshgrid(*[np.arange(s) for s in n], indexing='ij'), -1).reshape(-1, 3)
    m1 = mesh.reshape(-1)
    m2 = m1 if mesh2 is None else mesh2.reshape(-1)
    dk = 4 * np.pi / L.min() if dk is None else dk
    kmax = np.pi * (n / L).min()
    edges = np.arange(dk, kmax, dk) + dk / 2
    leg = {0: lambda u: 1.0, 1: lambda u: u, 2: lambda u: (3 * u * u - 1) / 2,
           4: lambda u: (35 * u ** 4 - 30 * u * u + 3) / 8}
    re = np.zeros((len(poles), len(edges) + 1))
    im = np.zeros_like(re)
    cnt = np.zeros(len(edges) + 1)
    ksum = np.zeros(len(edges) + 1)
    for idx in np.ndindex(*n):
        ph = np.exp(-2j * np.pi * (grid * (np.array(idx) / n)).sum(-1))
        f1 = (m1 * ph).sum() / np.sqrt(n.prod())
        f2 = (m2 * ph).sum() / np.sqrt(n.prod())
        kv = [2 * np.pi * (i if i <= (s - 1) // 2 else i - s) / l for i, s, l in zip(idx, n, L)]
        kk = float(np.sqrt(np.float32(sum(q * q for q in kv))))
        b = 0
        while b < len(edges) and edges[b] <= kk:
            b += 1
        mu = 0.0 if kk == 0 else sum(q * l for q, l in zip(kv, los)) / kk
        cnt[b] += 1
        ksum[b] += kk
        for i, ell in enumerate(poles):
            w = f1 * np.conj(f2) * (2 * ell + 1) * leg[ell](mu)
            re[i, b] += w.real
            im[i, b] += w.imag
    ps = re if mesh2 is None else np.hypot(re, im)
    return (ksum / cnt)[1:-1], (ps / cnt)[:, 1:-1] * np.prod(L / n)


@pytest.mark.parametrize("shape,box,dk", [((8, 8, 16), (32., 32., 64.), 0.15), ((6, 5, 7), (30., 25., 35.), 0.12),
                                          ((12, 12, 12), (12., 12., 12.), None)])
def test_power_spectrum_brute_force(shape, box, dk):
    """The default edges of a 12^3 unit-cell box sit at 3 and 5 times the
    fundamental, exactly on the |k| of modes (3,0,0), (3,4,0), (5,0,0)."""
    r = np.random.default_rng(2)
    mesh = r.normal(size=shape)
    k0, p0 = ou.power_spectrum(mesh, box_shape=box, kedges=dk)
    k1, p1 = brute_pk(mesh, box, dk=dk)
    assert len(k0) >= 1
    np.testing.assert_allclose(k0, k1, rtol=1e-12)
    np.testing.assert_allclose(p0, p1[0], rtol=1e-10)
    # cross spectra and multipoles, the line of sight normalised inside
    mesh2 = r.normal(size=shape)
    los = np.array([1., 2., 2.])
    k0, p0 = ou.power_spectrum(mesh, mesh2, box_shape=box, kedges=dk, multipoles=[0, 1, 2, 4], los=los)
    _, p1 = brute_pk(mesh, box, mesh2, (0, 1, 2, 4), los / 3.0, dk=dk)
    np.testing.assert_allclose(p0, p1, rtol=1e-10, atol=1e-12 * np.abs(p1).max())
    _, p0 = ou.power_spectrum(mesh, box_shape=box, kedges=dk, multipoles=[0, 2], los=los)
    _, p1 = brute_pk(mesh, box, None, (0, 2), los / 3.0, dk=dk)
    np.testing.assert_allclose(p0, p1, rtol=1e-10, atol=1e-12 * np.abs(p1).max())


def test_power_spectrum_multipoles_of_plane_waves():
    """A wave along the line of sight has mu = +-1 (P_ell = 1), one across
    it mu = 0 (P_2 = -1/2, P_4 = 3/8); mu divides by the float32 |k|, hence
    the 1e-6."""
    n, L = 16, 64.
    x = np.arange(n)
    wave = np.cos(2 * np.pi * 2 * x / n)
    along = np.broadcast_to(wave[None, None, :], (n, n, n))
    across = np.broadcast_to(wave[:, None, None], (n, n, n))
    _, pa = ou.power_spectrum(along, box_shape=(L,) * 3, multipoles=[0, 2, 4])
    _, pc = ou.power_spectrum(across, box_shape=(L,) * 3, multipoles=[0, 2, 4])
    b = np.argmax(pa[0])
    np.testing.assert_allclose(pa[:, b] / pa[0, b], [1, 5, 9], rtol=1e-6)
    np.testing.assert_allclose(pc[:, b] / pc[0, b], [1, -2.5, 9 * 3 / 8], rtol=1e-6)


def test_power_spectrum_cross_identities():
    r = np.random.default_rng(3)
    m = r.normal(size=(8, 8, 8))
    box = (40., 40., 40.)
    _, pa = ou.power_spectrum(m, box_shape=box)
    _, pc = ou.power_spectrum(m, m, box_shape=box)
    _, pn = ou.power_spectrum(m, -m, box_shape=box)     # |sum| of the complex bins
    np.testing.assert_allclose(pc, pa, rtol=1e-12)
    np.testing.assert_allclose(pn, pa, rtol=1e-12)
    _, t = ou.transfer(m, 3 * m, box)
    np.testing.assert_allclose(t, 3.0, rtol=1e-12)
    _, c = ou.coherence(m, m, box)
    np.testing.assert_allclose(c, 1.0, rtol=1e-12)


def test_power_spectrum_kedges_forms():
    m = np.random.default_rng(4).normal(size=(16, 16, 16))
    k, p = ou.power_spectrum(m, kedges=5)          # dk = kmax / 6
    assert len(k) == len(p) == len(np.arange(np.pi / 6, np.pi, np.pi / 6)) - 1
    k, p = ou.power_spectrum(m, kedges=0.3)
    assert len(k) == len(np.arange(0.3, np.pi, 0.3)) - 1
    edges = [0.2, 0.5, 1.0, 2.0]
    k, p = ou.power_spectrum(m, kedges=edges)
    assert len(k) == 3 and np.all((k > 0.2) & (k < 2.0))


def test_power_spectrum_plane_wave():
    n, L = 32, 100.
    m = 3
    x = np.arange(n)
    mesh = np.cos(2 * np.pi * m * x / n)[:, None, None] * np.ones((n, n, n))
    k, P = ou.power_spectrum(mesh, box_shape=(L, L, L))
    kw = 2 * np.pi * m / L
    b = np.argmin(abs(k - kw))
    # two modes of |FFT|^2 = (M/2)^2 / M each, over the bin's mode count
    nz = np.nonzero(P > 1e-12)[0]
    assert list(nz) == [b]
