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


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 12, 48)])
def test_force_vjp_matches_oracle(gpu, shape):
    """The sorted pipeline's force VJP (paint, divergence inside the R2C,
    potential, gather) against the oracle; (24, 12, 48) runs radix-3
    transforms."""
    from jaxpm_amd import integrate
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


SORTED_SHAPES = [(16, 16, 16), (8, 8, 8), (17, 33, 20), (32, 32, 64), (24, 12, 48)]


@pytest.mark.parametrize("shape", SORTED_SHAPES)
@pytest.mark.parametrize("relative", [False, True])
def test_sorted_force_kick_and_vjp(gpu, shape, relative):
    """The tile-sorted pipeline (potential + LDS stencil gathers) against the
    oracle's pm_forces and force VJP."""
    from jaxpm_amd import integrate
    from oracle import pm as opm
    print(f"sorted {shape} rel={relative}", flush=True)
    r = np.random.default_rng(35)
    n = int(np.prod(shape))
    disp = r.normal(scale=1.2, size=(n, 3)).astype(np.float32)
    pos = disp if relative else (lattice(shape) + disp).astype(np.float32)
    st = integrate.Stepper(shape, n, relative=relative)
    assert st.sorted
    xt = torch.as_tensor(pos, device="cuda")
    v = torch.full((n, 3), 0.25, device="cuda")
    phi = st.new_field()
    st.field_kick(xt, v, 0.5, phi)
    ref_f = opm.pm_forces(pos.reshape(*shape, 3) if relative else pos, mesh_shape=shape,
                          paint_absolute_pos=not relative).reshape(-1, 3)
    assert rel_err(v.cpu().numpy() - 0.25, 0.5 * ref_f) < 3e-5
    if relative:
        return  # the oracle adjoint restates the absolute painter
    u = r.normal(size=(n, 3)).astype(np.float32)
    g = torch.zeros_like(xt)
    st.vjp_field(phi, xt, torch.as_tensor(u, device="cuda"), 0.7, g)
    ref = 0.7 * oa.force_vjp(pos, u, shape)
    assert rel_err(g.cpu().numpy(), ref) < 5e-5
    # the same evaluation keeping its tile sort for the adjoint
    keep = st.new_field(keep_sort=True)
    st.field_kick(xt, torch.zeros_like(xt), 0.5, keep)
    g2 = torch.zeros_like(xt)
    st.vjp_field(keep, xt, torch.as_tensor(u, device="cuda"), 0.7, g2)
    assert rel_err(g2.cpu().numpy(), ref) < 5e-5


def test_library_kernel_timing(gpu):
    from jaxpm_amd import _lib, _ops
    _lib.timing_report()
    _lib.timing(True)
    m = torch.zeros((32, 32, 32), device="cuda")
    _ops.paint(m, torch.rand((4096, 3), device="cuda") * 32)
    torch.cuda.synchronize()
    _lib.timing(False)
    rep = _lib.timing_report()
    assert rep["bin_count_kernel"][0] == 1 and rep["bin_accum_kernel[nc=1]"][1] > 0
    assert _lib.timing_report() == {}


def test_sorted_force_kick_and_vjp_over_32768_tiles(gpu):
    """The tile-sorted force+kick and VJP on a mesh of 257 x 128 x 1 tiles
    (global-atomic counting sort)."""
    from jaxpm_amd import integrate
    from oracle import pm as opm
    shape = (4112, 2048, 1)
    print(f"sorted many tiles {shape}", flush=True)
    r = np.random.default_rng(45)
    n = 300000
    pos = (r.uniform(0, 1, size=(n, 3)) * np.array(shape)).astype(np.float32)
    st = integrate.Stepper(shape, n)
    assert st.sorted
    xt = torch.as_tensor(pos, device="cuda")
    v = torch.zeros((n, 3), device="cuda")
    field = st.new_field(keep_sort=True)
    st.field_kick(xt, v, 1.0, field)
    # float32 FFTs over 8.4M cells, with 1/k^2 amplifying the rounding of the
    # longest (4112-cell) modes: 1e-4 / 2e-4 rather than the 3e-5 / 5e-5 of
    # small meshes (the paint itself matches to 2e-6 at this shape)
    ref_f = opm.pm_forces(pos, mesh_shape=shape)
    assert rel_err(v.cpu().numpy(), ref_f) < 1e-4
    u = r.normal(size=(n, 3)).astype(np.float32)
    g = torch.zeros_like(xt)
    st.vjp_field(field, xt, torch.as_tensor(u, device="cuda"), 0.7, g)
    ref = 0.7 * oa.force_vjp(pos, u, shape)
    assert rel_err(g.cpu().numpy(), ref) < 2e-4


def test_sort_tuning_changes_speed_not_results(gpu, tmp_path, monkeypatch):
    """Every chunk / slice width computes the same forces and VJPs (float
    rounding), and the autotuner returns one of its candidates and caches it."""
    from jaxpm_amd import autotune, integrate
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    shape = (32, 32, 32)
    r = np.random.default_rng(46)
    n = int(np.prod(shape))
    pos = (lattice(shape) + r.normal(scale=2.0, size=(n, 3))).astype(np.float32)
    pos[:5000] = 10 + r.normal(scale=0.5, size=(5000, 3))      # a heavy tile: several chunks
    xt = torch.as_tensor(pos, device="cuda")
    u = torch.as_tensor(r.normal(size=(n, 3)).astype(np.float32), device="cuda")
    out = {}
    for tune in [(0, 0), (1, 1), (6, 7), (2, 5, 1), (0, 0, 1), (3, 2, 0, 1), (1, 1, 1, 1)]:
        print(f"tuning {tune}", flush=True)
        st = integrate.Stepper(shape, n, sort_tuning=tune)
        f = st.new_field(keep_sort=True)
        v = torch.zeros_like(xt)
        st.field_kick(xt, v, 1.0, f)
        g = torch.zeros_like(xt)
        st.vjp_field(f, xt, u, 0.5, g)
        out[tune] = (v.cpu().numpy(), g.cpu().numpy())
    for tune, (v, g) in out.items():
        assert rel_err(v, out[(0, 0)][0]) < 1e-6, tune
        assert rel_err(g, out[(0, 0)][1]) < 1e-6, tune
    best, timings = autotune.tune_stepper(shape, xt, reps=1)
    assert best in timings and len(timings) == len(autotune._candidates())
    again, t2 = autotune.tune_stepper(shape, xt)
    assert again == best and t2 == {}


@pytest.mark.parametrize("shape", [(32, 32, 32), (17, 33, 20)])
def test_tile_order_and_permutation(gpu, shape):
    """jpm_tile_order_u32 is a permutation that groups the particles by tile
    (tiles in increasing key order, keys from floor cells computed on the
    host), and
    jpm_permute_rows_f32's gather and scatter are inverse to each other."""
    from jaxpm_amd import _ops
    x0, _ = _state(shape, 40)
    x0[::7] += np.float32(shape[0])          # some positions a period away
    x0[::11] -= np.float32(shape[1])
    xt = torch.as_tensor(x0, device="cuda")
    order = _ops.tile_order(xt, shape).cpu().numpy().view(np.uint32)
    n = x0.shape[0]
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    cells = np.floor(x0.astype(np.float64)).astype(np.int64) % np.array(shape)
    t = [(s + 15) // 16 for s in shape]
    key = ((cells[:, 0] // 16) * t[1] + cells[:, 1] // 16) * t[2] + cells[:, 2] // 16
    assert np.all(np.diff(key[order]) >= 0)
    ot = torch.as_tensor(order.view(np.int32), device="cuda")
    g = _ops.permute_rows(ot, xt)
    assert np.array_equal(g.cpu().numpy(), x0[order])
    assert torch.equal(_ops.permute_rows(ot, g, scatter=True), xt)
    y = -2.0 * xt
    g2, h2 = _ops.permute_rows(ot, xt, y)
    assert torch.equal(g2, g)
    assert np.array_equal(h2.cpu().numpy(), y.cpu().numpy()[order])
    a, b = _ops.permute_rows(ot, g2, h2, scatter=True)
    assert torch.equal(a, xt) and torch.equal(b, y)


@pytest.mark.parametrize("shape", [(32, 32, 32), (17, 33, 20)])
def test_tile_ordered_sweeps_give_the_same_bits(gpu, shape):
    """NBody on particles in tile order (the default for absolute positions)
    against the caller's order, and with the sorts' tile counts taken from
    the kicks that drifted the positions against counted afresh: the same
    final state and gradients, bit for bit, since the paint sums in fixed
    point, the gathers are per particle and the counts are the same."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    x0, v0 = _state(shape, 41)
    r = np.random.default_rng(42)
    gx = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
    gv = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
    sched = integrate.schedule_tensor(integrate.kdk_schedule(pc.Cosmology(), np.linspace(0.1, 1.0, 6)))
    out = []
    # tile order with the kicks' tile counts chained into the next sorts,
    # tile order with every sort counting, the caller's order
    for reorder, chained in ((True, True), (True, False), (False, True)):
        st = integrate.Stepper(shape, x0.shape[0], reorder=reorder)
        assert st.reorder == reorder
        if not chained:
            st._tile_counts = None
        xt = torch.as_tensor(x0, device="cuda").requires_grad_(True)
        vt = torch.as_tensor(v0, device="cuda").requires_grad_(True)
        x, v = integrate.NBody.apply(xt, vt, sched, st, 2)
        ((x * gx).sum() + (v * gv).sum()).backward()
        out.append([t.detach().cpu().numpy() for t in (x, v, xt.grad, vt.grad)])
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 32, 64)])
def test_face_cells_apart_give_the_same_bits(gpu, shape):
    """The force evaluations' paint hands its tiles' low-face cells to the
    potential solve in a buffer of their own (absolute positions, meshes of
    full 16^3 tiles); JPM_TUNE_FACES_IN_MESH writes them into the mesh
    instead.  A leapfrog and its adjoint give the same bits either way, with
    a heavy tile split over several chunks (the reduce's general path)."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import integrate
    x0, v0 = _state(shape, 43)
    x0[:6000] = (10.0 + np.random.default_rng(44).normal(scale=0.8, size=(6000, 3))).astype(np.float32)
    r = np.random.default_rng(45)
    gx = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
    gv = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
    sched = integrate.schedule_tensor(integrate.kdk_schedule(pc.Cosmology(), np.linspace(0.1, 1.0, 6)))
    out = []
    for faces_in_mesh in (0, 1):
        print(f"shape {shape} faces_in_mesh {faces_in_mesh}", flush=True)
        st = integrate.Stepper(shape, x0.shape[0], sort_tuning=(0, 0, 0, 0, faces_in_mesh))
        xt = torch.as_tensor(x0, device="cuda").requires_grad_(True)
        vt = torch.as_tensor(v0, device="cuda").requires_grad_(True)
        x, v = integrate.NBody.apply(xt, vt, sched, st, 2)
        ((x * gx).sum() + (v * gv).sum()).backward()
        out.append([t.detach().cpu().numpy() for t in (x, v, xt.grad, vt.grad)])
    for a, b in zip(out[0], out[1]):
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


def test_kick_drift_velocity_out_and_field_only(gpu):
    """jpm_force_kick_drift_f32 with vel_out gives the in-place call's bits
    and leaves vel unchanged; jpm_force_kick_f32 with a null vel leaves the
    same potential as a full kick and no velocity change."""
    from jaxpm_amd import integrate
    shape = (32, 32, 32)
    x0, v0 = _state(shape, 36)
    n = x0.shape[0]
    st = integrate.Stepper(shape, n)
    x = torch.as_tensor(x0, device="cuda")
    v_in = torch.as_tensor(v0, device="cuda")
    # in place
    v_a, xo_a = v_in.clone(), torch.empty_like(x)
    f_a = st.new_field(keep_sort=True)
    st.field_kick_drift(x, v_a, 0.3, 0.2, f_a, xo_a)
    # into vel_out
    v_b, vo_b, xo_b = v_in.clone(), torch.empty_like(x), torch.empty_like(x)
    f_b = st.new_field(keep_sort=True)
    st.field_kick_drift(x, v_b, 0.3, 0.2, f_b, xo_b, vout=vo_b)
    torch.cuda.synchronize()
    assert torch.equal(vo_b, v_a) and torch.equal(xo_b, xo_a)
    assert torch.equal(v_b, v_in)
    assert torch.equal(f_a[0], f_b[0])
    # the force evaluation alone: same potential, no kick
    f_c = st.new_field(keep_sort=True)
    st.field_kick(x, None, 0.3, f_c)
    v_d = v_in.clone()
    f_d = st.new_field(keep_sort=True)
    st.field_kick(x, v_d, 0.3, f_d)
    torch.cuda.synchronize()
    assert torch.equal(f_c[0], f_d[0])
    assert not torch.equal(v_d, v_in)
