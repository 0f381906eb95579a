"""GPU parity of CIC paint / read and their adjoints against the oracle."""
import numpy as np
import pytest
import torch

from oracle import painting as op

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.as_tensor(np.asarray(x), dtype=torch.float32, device="cuda")


def host(t):
    return t.detach().cpu().numpy().astype(np.float64)


def lattice(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).astype(np.float32)


def close(got, ref, rel=2e-6):
    """FP32 sums of O(10) deposits: error relative to the largest value."""
    scale = max(np.abs(ref).max(), 1.0)
    err = np.abs(got - ref).max() / scale
    assert err < rel, f"max error {err:.3e} (x{scale:.3g})"


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (32, 32, 64)])
def test_paint_absolute(gpu, shape):
    from jaxpm_amd import painting
    print(f"paint absolute {shape}", flush=True)
    r = np.random.default_rng(10)
    n = int(np.prod(shape))
    L = np.array(shape)
    pos = r.uniform(-0.5, 1.5, size=(n, 3)).astype(np.float32) * L  # wrap on both sides
    w = r.normal(size=n).astype(np.float32)
    base = r.normal(size=shape).astype(np.float32)
    got = host(painting.cic_paint(dev(base), dev(pos)))
    close(got, op.cic_paint(base, pos))
    got = host(painting.cic_paint(dev(np.zeros(shape)), dev(pos), weight=dev(w)))
    close(got, op.cic_paint(np.zeros(shape), pos, w))


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (32, 32, 64)])
def test_paint_relative(gpu, shape):
    from jaxpm_amd import painting
    print(f"paint relative {shape}", flush=True)
    r = np.random.default_rng(11)
    disp = r.normal(scale=3.0, size=(*shape, 3)).astype(np.float32)
    disp[0, 0, 0] = (-1e-9, 0.0, 0.0)   # rounds onto index L in float32: dropped
    got = host(painting.cic_paint_dx(dev(disp)))
    ref = op.cic_paint_dx(disp, dtype=np.float32).astype(np.float64)
    close(got, ref)
    w = r.uniform(0.5, 2.0, size=shape).astype(np.float32)
    close(host(painting.cic_paint_dx(dev(disp), weight=dev(w))), op.cic_paint_dx(disp, w, dtype=np.float32))


def test_paint_weight_shape_error(gpu):
    from jaxpm_amd import painting
    with pytest.raises(ValueError, match="Weight shape must match particle shape"):
        painting.cic_paint_dx(dev(np.zeros((4, 4, 4, 3))), weight=dev(np.ones(5)))


def test_deterministic_paint_is_bitwise_reproducible(gpu):
    from jaxpm_amd import painting
    shape = (32, 32, 32)
    r = np.random.default_rng(12)
    # clustered: every particle in a few cells -> heavy atomic contention
    pos = (16 + r.normal(scale=1.5, size=(32768, 3))).astype(np.float32)
    a = painting.cic_paint(dev(np.zeros(shape)), dev(pos), deterministic=True)
    b = painting.cic_paint(dev(np.zeros(shape)), dev(pos), deterministic=True)
    assert torch.equal(a, b)
    plain = painting.cic_paint(dev(np.zeros(shape)), dev(pos))
    close(host(a), host(plain), rel=1e-5)  # float32 atomics lose ~1e-7 per deposit
    close(host(a), op.cic_paint(np.zeros(shape), pos), rel=5e-7)


def test_paint_large_mass_conservation(gpu):
    """256^3 particles on a 256^3 mesh: sum(mesh) == N, uniform and clustered."""
    from jaxpm_amd import painting
    shape = (256, 256, 256)
    n = 256 ** 3
    g = torch.Generator(device="cuda").manual_seed(1)
    pos = torch.rand((n, 3), generator=g, device="cuda") * 256
    for p in (pos, 128 + 4 * torch.randn((n, 3), generator=g, device="cuda")):
        m = painting.cic_paint(torch.zeros(shape, device="cuda"), p)
        total = m.double().sum().item()
        assert abs(total - n) < 1e-5 * n, total
        md = painting.cic_paint(torch.zeros(shape, device="cuda"), p, deterministic=True)
        assert abs(md.double().sum().item() - n) < 1e-6 * n
        assert (m - md).abs().max().item() < 1e-3 * max(1.0, md.abs().max().item())


def test_paint_128_vs_oracle(gpu):
    from jaxpm_amd import painting
    shape = (128, 128, 128)
    r = np.random.default_rng(13)
    pos = (lattice(shape).reshape(-1, 3) + r.normal(scale=2.0, size=(128 ** 3, 3))).astype(np.float32)
    close(host(painting.cic_paint(dev(np.zeros(shape)), dev(pos))), op.cic_paint(np.zeros(shape), pos))


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (32, 32, 64)])
def test_read_absolute_and_relative(gpu, shape):
    from jaxpm_amd import painting
    print(f"read {shape}", flush=True)
    r = np.random.default_rng(14)
    mesh = r.normal(size=shape).astype(np.float32)
    n = int(np.prod(shape))
    pos = (r.uniform(-0.5, 1.5, size=(n, 3)) * np.array(shape)).astype(np.float32)
    close(host(painting.cic_read(dev(mesh), dev(pos))), op.cic_read(mesh, pos))
    disp = r.normal(scale=3.0, size=(*shape, 3)).astype(np.float32)
    close(host(painting.cic_read_dx(dev(mesh), dev(disp))), op.cic_read_dx(mesh, disp, dtype=np.float32))
    # lattice read returns node values exactly
    np.testing.assert_array_equal(host(painting.cic_read(dev(mesh), dev(lattice(shape)))), mesh)
    # three meshes in one call
    m3 = r.normal(size=(3, *shape)).astype(np.float32)
    got = host(painting.cic_read(dev(m3), dev(pos)))
    for c in range(3):
        close(got[:, c], op.cic_read(m3[c], pos))


def test_read_paint_transpose_and_position_gradients(gpu):
    from jaxpm_amd import painting
    shape = (16, 16, 16)
    r = np.random.default_rng(15)
    n = 4096
    pos = r.uniform(0, 16, size=(n, 3))
    # keep away from cell faces so that central differences stay on one piece
    frac = pos - np.floor(pos)
    pos = np.floor(pos) + np.clip(frac, 0.01, 0.99)
    pos = pos.astype(np.float32)
    g = r.normal(size=shape).astype(np.float32)
    w = r.normal(size=n).astype(np.float32)
    lhs = (painting.cic_paint(dev(np.zeros(shape)), dev(pos), weight=dev(w)).double() * dev(g).double()).sum()
    rhs = (dev(w).double() * painting.cic_read(dev(g), dev(pos)).double()).sum()
    assert abs(lhs.item() - rhs.item()) < 1e-5 * abs(lhs.item())

    # d/dpos of sum(u * read(g, pos)) against float64 central differences
    u = r.normal(size=n)
    pt = dev(pos).requires_grad_(True)
    (painting.cic_read(dev(g), pt) * dev(u)).sum().backward()
    eps = 1e-4
    dread = np.zeros((n, 3))
    for d in range(3):
        e = np.zeros(3)
        e[d] = eps
        dread[:, d] = (op.cic_read(g, pos + e) - op.cic_read(g, pos - e)) / (2 * eps)
    close(host(pt.grad), dread * u[:, None], rel=1e-4)

    # d/dpos of <paint(pos, w), g>: same derivative with u = w
    pt = dev(pos).requires_grad_(True)
    (painting.cic_paint(dev(np.zeros(shape)), pt, weight=dev(w)) * dev(g)).sum().backward()
    close(host(pt.grad), dread * w[:, None], rel=1e-4)

    # mesh cotangent of read is the paint of the cotangent
    mt = dev(g).requires_grad_(True)
    (painting.cic_read(mt, dev(pos)) * dev(u)).sum().backward()
    close(host(mt.grad), op.cic_paint(np.zeros(shape), pos, u))


def test_compensate_cic(gpu):
    from jaxpm_amd import painting
    r = np.random.default_rng(16)
    f = r.normal(size=(16, 16, 32)).astype(np.float32)
    close(host(painting.compensate_cic(dev(f))), op.compensate_cic(f), rel=5e-6)


# ----------------------------------------------------------- binned paint

BINNED_SHAPES = [(8, 8, 8), (12, 10, 9), (17, 33, 20), (48, 48, 48), (32, 32, 64)]


@pytest.mark.parametrize("shape", BINNED_SHAPES)
@pytest.mark.parametrize("deterministic", [False, True])
def test_binned_paint_absolute(gpu, shape, deterministic):
    from jaxpm_amd import _ops
    print(f"binned abs {shape} det={deterministic}", flush=True)
    r = np.random.default_rng(40)
    n = int(np.prod(shape))
    pos = (r.uniform(-0.5, 1.5, size=(n, 3)) * np.array(shape)).astype(np.float32)
    w = r.normal(size=n).astype(np.float32)
    base = r.normal(size=shape).astype(np.float32)
    m = dev(base)
    _ops.paint(m, dev(pos), deterministic=deterministic)                       # accumulate
    close(host(m), op.cic_paint(base, pos))
    m = dev(np.zeros(shape))
    _ops.paint(m, dev(pos), weight=dev(w), deterministic=deterministic, zero=True)
    close(host(m), op.cic_paint(np.zeros(shape), pos, w))
    # weights not 16-byte aligned (the weight bound's vector loads start
    # after a scalar head)
    wpad = dev(np.concatenate([[0.0], w]).astype(np.float32))
    m = dev(np.zeros(shape))
    _ops.paint(m, dev(pos), weight=wpad[1:], deterministic=deterministic, zero=True)
    close(host(m), op.cic_paint(np.zeros(shape), pos, w))


@pytest.mark.parametrize("shape", BINNED_SHAPES)
def test_binned_paint_relative(gpu, shape):
    from jaxpm_amd import _ops
    r = np.random.default_rng(41)
    disp = r.normal(scale=3.0, size=(*shape, 3)).astype(np.float32)
    disp[0, 0, 0] = (-1e-9, 0.0, 0.0)
    m = dev(np.zeros(shape))
    _ops.paint(m, dev(disp), relative=True, zero=True)
    close(host(m), op.cic_paint_dx(disp, dtype=np.float32))
    m2 = dev(np.zeros(shape))
    _ops.paint(m2, dev(disp), relative=True, zero=True, deterministic=True)
    close(host(m2), op.cic_paint_dx(disp, dtype=np.float32))


@pytest.mark.parametrize("deterministic", [False, True])
def test_binned_paint_three_channels(gpu, deterministic):
    from jaxpm_amd import _ops
    shape = (24, 40, 32)
    r = np.random.default_rng(42)
    n = 30000
    pos = (r.uniform(0, 1, size=(n, 3)) * np.array(shape)).astype(np.float32)
    w3 = r.normal(size=(n, 3)).astype(np.float32)
    got = host(_ops.paint_multi(shape, dev(pos), dev(w3), deterministic=deterministic))
    for c in range(3):
        close(got[c], op.cic_paint(np.zeros(shape), pos, w3[:, c]))
    atomic = host(_ops.paint_multi(shape, dev(pos), dev(w3), method="atomic"))
    close(got, atomic, rel=1e-5)


def test_binned_paint_heavy_tile_and_empty(gpu):
    """One tile holding 50k particles (13 chunks), most tiles empty, and the
    bitwise reproducibility of the fixed-point mode."""
    from jaxpm_amd import _ops
    shape = (64, 64, 64)
    r = np.random.default_rng(43)
    pos = (20 + r.uniform(0, 8, size=(50000, 3))).astype(np.float32)
    pos[:10] = r.uniform(0, 64, size=(10, 3))
    m = dev(np.zeros(shape))
    _ops.paint(m, dev(pos), zero=True)
    close(host(m), op.cic_paint(np.zeros(shape), pos), rel=2e-6)
    a = dev(np.zeros(shape))
    b = dev(np.zeros(shape))
    _ops.paint(a, dev(pos), zero=True, deterministic=True)
    _ops.paint(b, dev(pos), zero=True, deterministic=True)
    assert torch.equal(a, b)
    close(host(a), op.cic_paint(np.zeros(shape), pos), rel=5e-7)
    # no particles: zero mesh / untouched mesh
    z = dev(np.ones(shape))
    _ops.paint(z, dev(np.zeros((0, 3))), zero=True)
    assert float(z.abs().max()) == 0.0


def test_binned_matches_atomic_at_256(gpu):
    from jaxpm_amd import _ops
    shape = (256, 256, 256)
    n = 256 ** 3
    g = torch.Generator(device="cuda").manual_seed(3)
    for p in (torch.rand((n, 3), generator=g, device="cuda") * 256,
              128 + 6 * torch.randn((n, 3), generator=g, device="cuda")):
        a = torch.zeros(shape, device="cuda")
        b = torch.zeros(shape, device="cuda")
        _ops.paint(a, p, method="binned")
        _ops.paint(b, p, method="atomic")
        scale = max(1.0, float(b.abs().max()))
        assert float((a - b).abs().max()) < 2e-5 * scale
        assert abs(a.double().sum().item() - n) < 1e-5 * n


# Over 32768 tiles the tile sort counts and scatters with wavefront-run
# global atomics instead of LDS histograms; a thin mesh reaches that regime
# with few cells (257 x 128 x 1 tiles).
MANY_TILES = (4112, 2048, 1)


@pytest.mark.parametrize("deterministic", [False, True])
def test_binned_paint_over_32768_tiles(gpu, deterministic):
    from jaxpm_amd import _ops
    print(f"binned many tiles det={deterministic}", flush=True)
    shape = MANY_TILES
    r = np.random.default_rng(44)
    n = 400000
    pos = (r.uniform(-0.2, 1.2, size=(n, 3)) * np.array(shape)).astype(np.float32)
    pos[:50000] = (r.normal(size=(50000, 3)) * 3 + np.array([100.0, 300.0, 0.5])).astype(np.float32)
    m = dev(np.zeros(shape))
    _ops.paint(m, dev(pos), zero=True, deterministic=deterministic)
    close(host(m), op.cic_paint(np.zeros(shape), pos), rel=2e-6)
    w3 = r.normal(size=(n, 3)).astype(np.float32)
    got = host(_ops.paint_multi(shape, dev(pos), dev(w3), deterministic=deterministic))
    for c in range(3):
        close(got[c], op.cic_paint(np.zeros(shape), pos, w3[:, c]), rel=2e-6)


@pytest.mark.parametrize("shape", [(16, 16, 16), (17, 33, 20), (64, 32, 48)])
@pytest.mark.parametrize("nc", [1, 3])
def test_sorted_read_matches_direct(gpu, shape, nc):
    """The tile-sorted read gives the direct read's values bit for bit
    (positions wrapping on both sides, a heavy tile split over chunks), and
    matches the oracle."""
    from jaxpm_amd import _ops, painting
    r = np.random.default_rng(60)
    n = 3 * int(np.prod(shape))
    pos = (r.uniform(-0.5, 1.5, size=(n, 3)) * np.array(shape)).astype(np.float32)
    heavy = min(20000, n // 2)
    pos[:heavy] = 3.3 + r.normal(scale=0.4, size=(heavy, 3))
    meshes = r.normal(size=(nc, *shape)).astype(np.float32)
    m = dev(meshes if nc > 1 else meshes[0])
    direct = host(_ops.read(m, dev(pos)))
    srt = host(_ops.read(m, dev(pos), method="sorted"))
    np.testing.assert_array_equal(srt, direct)
    ref = np.stack([op.cic_read(meshes[c], pos) for c in range(nc)], -1)
    close(srt.reshape(n, nc), ref)
    api = host(painting.cic_read(dev(meshes[0]), dev(pos), method="sorted"))
    np.testing.assert_array_equal(api, host(_ops.read(dev(meshes[0]), dev(pos))))
