"""Batched leapfrog (one launch per stage over simulations of different
meshes) against the single-simulation path, simulation by simulation."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (16, 16, 16), (24, 16, 32), (8, 12, 8)]


def lattice(shape):
    return np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)


def states(shapes, relative, seed):
    r = np.random.default_rng(seed)
    xs, vs = [], []
    for sh in shapes:
        n = int(np.prod(sh))
        d = r.normal(scale=0.8, size=(n, 3))
        xs.append((d if relative else lattice(sh) + d).astype(np.float32))
        vs.append(r.normal(scale=0.3, size=(n, 3)).astype(np.float32))
    return xs, vs


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("relative", [False, True])
def test_batch_matches_single(gpu, relative):
    from jaxpm_amd import integrate
    print(f"batch relative={relative}", flush=True)
    xs, vs = states(SHAPES, relative, 50)
    r = np.random.default_rng(51)
    S = 4
    coeffs = np.concatenate([r.uniform(0.02, 0.06, (len(SHAPES), 1)), r.uniform(0.1, 0.3, (len(SHAPES), S)),
                             r.uniform(0.03, 0.08, (len(SHAPES), S))], 1)
    gxs = [r.normal(size=x.shape) for x in xs]
    gvs = [r.normal(size=x.shape) for x in xs]
    # batched
    st = integrate.BatchStepper(SHAPES, [x.shape[0] for x in xs], relative=relative)
    xt = torch.as_tensor(np.concatenate(xs), device="cuda").requires_grad_(True)
    vt = torch.as_tensor(np.concatenate(vs), device="cuda").requires_grad_(True)
    ct = torch.tensor(coeffs, dtype=torch.float64, requires_grad=True)
    xb, vb = integrate.BatchNBody.apply(xt, vt, ct, st, 2)
    gx = torch.as_tensor(np.concatenate(gxs), device="cuda", dtype=torch.float32)
    gv = torch.as_tensor(np.concatenate(gvs), device="cuda", dtype=torch.float32)
    ((xb * gx).sum() + (vb * gv).sum()).backward()
    xb, vb = xb.detach().cpu().numpy(), vb.detach().cpu().numpy()
    gxb, gvb, gcb = xt.grad.cpu().numpy(), vt.grad.cpu().numpy(), ct.grad.numpy()
    # one simulation at a time
    off = 0
    for b, sh in enumerate(SHAPES):
        n = xs[b].shape[0]
        s1 = integrate.Stepper(sh, n, relative=relative)
        x1 = torch.as_tensor(xs[b], device="cuda").requires_grad_(True)
        v1 = torch.as_tensor(vs[b], device="cuda").requires_grad_(True)
        c1 = torch.tensor(coeffs[b], dtype=torch.float64, requires_grad=True)
        xo, vo = integrate.NBody.apply(x1, v1, c1, s1, 2)
        ((xo * gx[off:off + n]).sum() + (vo * gv[off:off + n]).sum()).backward()
        sl = slice(off, off + n)
        assert rel_err(xb[sl] - xs[b], xo.detach().cpu().numpy() - xs[b]) < 1e-5, b
        assert rel_err(vb[sl], vo.detach().cpu().numpy()) < 1e-5, b
        assert rel_err(gxb[sl], x1.grad.cpu().numpy()) < 1e-4, b
        assert rel_err(gvb[sl], v1.grad.cpu().numpy()) < 1e-4, b
        np.testing.assert_allclose(gcb[b], c1.grad.numpy(), rtol=1e-4, atol=1e-4 * np.abs(c1.grad.numpy()).max())
        off += n


def test_batch_vectors(gpu):
    from jaxpm_amd import integrate
    shapes = [(8, 8, 8), (8, 8, 8), (12, 8, 8)]
    counts = [100, 512, 37]
    st = integrate.BatchStepper(shapes, counts)
    r = np.random.default_rng(52)
    x = r.normal(size=(sum(counts), 3)).astype(np.float32)
    y = r.normal(size=(sum(counts), 3)).astype(np.float32)
    a = np.array([0.5, -1.5, 2.0], np.float32)
    yt = torch.as_tensor(y, device="cuda")
    st.axpy(torch.as_tensor(a, device="cuda"), torch.as_tensor(x, device="cuda"), yt)
    ref = y.copy()
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    st.dot_into(out, torch.as_tensor(x, device="cuda"), torch.as_tensor(y, device="cuda"))
    off = 0
    for b, n in enumerate(counts):
        ref[off:off + n] += a[b] * x[off:off + n]
        d = np.sum(x[off:off + n].astype(np.float64) * y[off:off + n].astype(np.float64))
        assert abs(out[b].item() - d) < 1e-12 * np.sum(np.abs(x[off:off + n] * y[off:off + n]))
        off += n
    np.testing.assert_array_equal(yt.cpu().numpy(), ref)


def test_batch_rejects_interleaved_shapes(gpu):
    from jaxpm_amd import integrate
    with pytest.raises(ValueError, match="contiguous"):
        integrate.BatchStepper([(8, 8, 8), (16, 16, 16), (8, 8, 8)], [512, 4096, 512])


def test_batch_graph_replays_match_eager(gpu):
    """The captured time loop, replayed several times with new inputs and
    coefficients and eager work in between, against eager runs."""
    from jaxpm_amd import integrate
    print("batch graph replays", flush=True)
    xs, vs = states(SHAPES, False, 53)
    r = np.random.default_rng(54)
    S = 3
    n = [x.shape[0] for x in xs]

    def coeffs():
        return np.concatenate([r.uniform(0.02, 0.06, (len(SHAPES), 1)), r.uniform(0.1, 0.3, (len(SHAPES), S)),
                               r.uniform(0.03, 0.08, (len(SHAPES), S))], 1)

    st = integrate.BatchStepper(SHAPES, n)
    c0 = coeffs()
    eager = integrate.BatchLeapfrog(st, c0, 2, coeff_grads=True, graph=False)
    graphed = integrate.BatchLeapfrog(st, c0, 2, coeff_grads=True, graph=True)
    x0 = torch.as_tensor(np.concatenate(xs), device="cuda")
    v0 = torch.as_tensor(np.concatenate(vs), device="cuda")
    for rep in range(3):
        c = coeffs()
        eager.sc.update(c)
        graphed.sc.update(c)
        dx = torch.as_tensor(r.normal(scale=0.05, size=x0.shape).astype(np.float32), device="cuda")
        gx = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
        gv = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
        xg, vg = [t.clone() for t in graphed.forward(x0 + dx, v0)]
        lg = [t.clone() for t in graphed.backward(gx, gv)]
        torch.cuda.synchronize()
        xe, ve = [t.clone() for t in eager.forward(x0 + dx, v0)]
        le = [t.clone() for t in eager.backward(gx, gv)]
        torch.cuda.synchronize()
        print(f"replay {rep}: bitwise x={torch.equal(xg, xe)} lx={torch.equal(lg[0], le[0])}", flush=True)
        for a, b in zip([xg, vg] + lg, [xe, ve] + le):
            scale = float(b.abs().max())
            assert float((a - b).abs().max()) <= 1e-6 * scale, rep


def test_fully_batched_model_matches_per_simulation(gpu):
    """linear_field -> 1LPT -> leapfrog -> paint, every stage batched, against
    the per-simulation model (same noise, cosmologies, steps): densities and
    noise gradients."""
    from jaxpm_amd import cosmology as pc
    from jaxpm_amd import diffsim
    print("fully batched model", flush=True)
    shapes = [(16, 16, 16), (16, 16, 16), (24, 24, 24), (32, 32, 32)]
    boxes = [(4.0 * s[0],) * 3 for s in shapes]
    cos = [pc.Cosmology(Omega_c=oc, Omega_b=0.05, h=0.7, n_s=0.96, sigma8=s8)
           for oc, s8 in [(0.25, 0.8), (0.3, 0.75), (0.22, 0.85), (0.33, 0.72)]]
    a_steps = np.linspace(0.1, 1.0, 6)
    r = np.random.default_rng(55)
    noises = [r.normal(size=sh).astype(np.float32) for sh in shapes]
    targets = [r.normal(scale=0.3, size=sh).astype(np.float32) for sh in shapes]
    model = diffsim.BatchLptNBody(shapes, boxes, a_steps, cos, order=1, checkpoint_every=2, graph=True)
    nt = torch.as_tensor(np.concatenate([n.reshape(-1) for n in noises]), device="cuda").requires_grad_(True)
    tt = torch.as_tensor(np.concatenate([t.reshape(-1) for t in targets]), device="cuda")
    rho = model.forward_fixed_cosmology(nt)
    ((rho - 1.0 - tt) ** 2).sum().backward()
    rho, g = rho.detach().cpu().numpy(), nt.grad.cpu().numpy()
    off = 0
    for b, sh in enumerate(shapes):
        m = int(np.prod(sh))
        single = diffsim.LptNBody(sh, boxes[b], a_steps, base_cosmo=cos[b], order=1, checkpoint_every=2)
        n1 = torch.as_tensor(noises[b], device="cuda").requires_grad_(True)
        r1 = single(n1, single.parameter_vector())
        ((r1 - 1.0 - torch.as_tensor(targets[b], device="cuda")) ** 2).sum().backward()
        assert rel_err(rho[off:off + m], r1.detach().cpu().numpy().reshape(-1)) < 1e-5, b
        assert rel_err(g[off:off + m], n1.grad.cpu().numpy().reshape(-1)) < 1e-4, b
        off += m


def test_batch_tuning_changes_speed_not_results(gpu):
    """Chunk, slice, density-window and XCD-placement choices give the same
    batched kick and force VJP up to float rounding."""
    from jaxpm_amd import integrate
    xs, vs = states(SHAPES, False, 56)
    r = np.random.default_rng(57)
    n = [x.shape[0] for x in xs]
    x = torch.as_tensor(np.concatenate(xs), device="cuda")
    u = torch.as_tensor(r.normal(size=x.shape).astype(np.float32), device="cuda")
    out = {}
    for tune in [(0, 0), (1, 2, 0, 1), (2, 1, 1, 1), (0, 0, 1, 0)]:
        st = integrate.BatchStepper(SHAPES, n, sort_tuning=tune)
        f = st.new_field(keep_sort=True)
        gam = torch.full((st.B,), 0.7, dtype=torch.float32, device="cuda")
        one = torch.ones(st.B, dtype=torch.float32, device="cuda")
        v = torch.zeros_like(x)
        st.field_kick(x, v, gam, f)
        g = torch.zeros_like(x)
        st.vjp_field(f, x, u, one, one, g)
        out[tune] = (v.cpu().numpy(), g.cpu().numpy())
    for tune, (v, g) in out.items():
        assert rel_err(v, out[(0, 0)][0]) < 1e-6, tune
        assert rel_err(g, out[(0, 0)][1]) < 1e-6, tune


@pytest.mark.parametrize("graph", [False, True])
def test_fused_kick_drift_matches_separate_drift(gpu, graph):
    """The kick-drift fused sweeps (no coefficient cotangents) against the
    sweeps with separate drifts (coefficient cotangents on): the same float
    operations, so states and adjoints agree bit for bit."""
    from jaxpm_amd import integrate
    xs, vs = states(SHAPES, False, 58)
    r = np.random.default_rng(59)
    S = 5
    c = np.concatenate([r.uniform(0.02, 0.06, (len(SHAPES), 1)), r.uniform(0.1, 0.3, (len(SHAPES), S)),
                        r.uniform(0.03, 0.08, (len(SHAPES), S))], 1)
    st = integrate.BatchStepper(SHAPES, [x.shape[0] for x in xs])
    x0 = torch.as_tensor(np.concatenate(xs), device="cuda")
    v0 = torch.as_tensor(np.concatenate(vs), device="cuda")
    gx = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
    gv = torch.as_tensor(r.normal(size=x0.shape).astype(np.float32), device="cuda")
    out = []
    for want_c in (False, True):
        lf = integrate.BatchLeapfrog(st, c, 2, coeff_grads=want_c, graph=graph)
        # the fused sweeps run on the particles in tile order, the others in
        # the caller's order: the same bits either way
        assert (lf.order is not None) == (not want_c)
        x, v = [t.clone() for t in lf.forward(x0, v0)]
        lx, lv = [t.clone() for t in lf.backward(gx, gv)[:2]]
        out.append((x, v, lx, lv))
    for a, b in zip(*out):
        assert torch.equal(a, b)


def test_batch_tile_order(gpu):
    """jpm_batch_tile_order_u32: a permutation that keeps each simulation's
    slots within its own particle range and groups them by tile."""
    from jaxpm_amd import integrate
    xs, _ = states(SHAPES, False, 60)
    st = integrate.BatchStepper(SHAPES, [x.shape[0] for x in xs])
    x = torch.as_tensor(np.concatenate(xs), device="cuda")
    order = st.tile_order(x, torch.empty(st.N, dtype=torch.int32, device="cuda")).cpu().numpy()
    for b, sh in enumerate(SHAPES):
        p0, n = int(st.part0[b]), st.counts[b]
        o = order[p0:p0 + n]
        assert np.array_equal(np.sort(o), np.arange(p0, p0 + n))
        cells = np.floor(xs[b].astype(np.float64)).astype(np.int64) % np.array(sh)
        t = [(s + 15) // 16 for s in sh]
        key = ((cells[:, 0] // 16) * t[1] + cells[:, 1] // 16) * t[2] + cells[:, 2] // 16
        assert np.all(np.diff(key[o - p0]) >= 0)
