"""The leapfrog drivers' buffer handling, on the CPU.

`integrate.NBody` and `integrate.BatchLeapfrog` keep checkpoints and each
step's positions in the buffers the fused kick-drifts write, recompute
segments in place on their checkpoints and permute into and out of the
tile order.  These tests drive them with a stand-in stepper whose force is
a smooth elementwise function: every force VJP must be handed exactly the
positions its force evaluation saw, and the gradients must equal autograd
through the same leapfrog written as tensor expressions.
"""
import numpy as np
import pytest
import torch

from jaxpm_amd import integrate


def force(x):
    return -torch.sin(x)


def force_vjp(x, cot):
    return -torch.cos(x) * cot


class Field(list):
    """Stand-in for a force evaluation's kept data: the positions it saw."""


class MockStepper:
    """The stepper interface NBody drives, in float64 torch on the CPU."""

    def __init__(self, reorder):
        self.reorder = reorder
        self.vjps = 0
        self.field_only = 0

    def tile_order(self, x):
        return torch.argsort(x[:, 0] + 0.37 * x[:, 1]).to(torch.int32)

    def permute(self, order, *arrays, scatter=False):
        idx = order.long()
        outs = []
        for a in arrays:
            if scatter:
                o = torch.empty_like(a)
                o[idx] = a
            else:
                o = a[idx].clone()
            outs.append(o)
        return tuple(outs) if len(outs) > 1 else outs[0]

    def new_field(self, keep_sort=False):
        return Field([None])

    def field_kick(self, x, v, gamma, field, chain=False):
        field[0] = x.clone()
        if v is None:
            self.field_only += 1
        else:
            v += float(gamma) * force(x)

    def counts_like(self):
        return None

    def counts_save(self, out):
        pass

    def counts_load(self, src):
        pass

    def field_kick_drift(self, x, v, gamma, beta, field, xout, xkeep=None, chain=False, vout=None):
        if vout is not None:
            assert vout is not v
            vout.copy_(v)
            v = vout
        self.field_kick(x, v, gamma, field)
        y = x + float(beta) * v
        xout.copy_(y)
        if xkeep is not None:
            xkeep.copy_(y)

    def _vjp(self, field, x, cot, scale, gpos):
        assert torch.equal(field[0], x), "force VJP handed other positions than its evaluation saw"
        gpos += scale * force_vjp(x, cot)
        self.vjps += 1

    def vjp_field(self, field, x, cot, scale, gpos):
        self._vjp(field, x, cot, float(scale), gpos)

    def vjp_field_drift(self, field, x, cot, scale, beta, gpos, chain=False):
        self._vjp(field, x, cot, float(scale), gpos)
        cot += float(beta) * gpos


def plain_leapfrog(x0, v0, k0, beta, gamma):
    x, v = x0, v0 + k0 * force(x0)
    for b, g in zip(beta, gamma):
        x = x + float(b) * v
        v = v + float(g) * force(x)
    return x, v


def schedule(S, seed):
    r = np.random.default_rng(seed)
    return r.uniform(0.05, 0.2), r.uniform(0.2, 0.5, S), r.uniform(0.1, 0.3, S)


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("S,every", [(1, None), (4, None), (7, 3), (6, 2), (5, 1)])
def test_nbody_buffers_and_adjoint(reorder, S, every):
    k0, beta, gamma = schedule(S, 10 + S)
    r = np.random.default_rng(S)
    x0 = torch.as_tensor(r.uniform(-2, 2, (40, 3)))
    v0 = torch.as_tensor(r.normal(size=(40, 3)))
    gx = torch.as_tensor(r.normal(size=(40, 3)))
    gv = torch.as_tensor(r.normal(size=(40, 3)))
    coeffs = torch.as_tensor(np.concatenate([[k0], beta, gamma]))
    st = MockStepper(reorder)
    a, b = x0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    x, v = integrate.NBody.apply(a, b, coeffs, st, every)
    ((x * gx).sum() + (v * gv).sum()).backward()
    assert st.vjps == S + 1
    # one force evaluation without its kick per recomputed segment
    k = S if every is None else every
    assert st.field_only == len(range(0, S, k)) - 1
    ra, rb = x0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    xr, vr = plain_leapfrog(ra, rb, k0, beta, gamma)
    ((xr * gx).sum() + (vr * gv).sum()).backward()
    torch.testing.assert_close(x.detach(), xr.detach(), rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(v.detach(), vr.detach(), rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(a.grad, ra.grad, rtol=1e-10, atol=1e-10)
    torch.testing.assert_close(b.grad, rb.grad, rtol=1e-10, atol=1e-10)


def test_nbody_forward_without_grad():
    k0, beta, gamma = schedule(6, 3)
    r = np.random.default_rng(3)
    x0 = torch.as_tensor(r.uniform(-2, 2, (30, 3)))
    v0 = torch.as_tensor(r.normal(size=(30, 3)))
    coeffs = torch.as_tensor(np.concatenate([[k0], beta, gamma]))
    x, v = integrate.NBody.apply(x0, v0, coeffs, MockStepper(True), 2)
    xr, vr = plain_leapfrog(x0, v0, k0, beta, gamma)
    torch.testing.assert_close(x, xr, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(v, vr, rtol=1e-12, atol=1e-12)


class MockBatchStepper(MockStepper):
    """The batched stepper interface BatchLeapfrog drives: coefficients are
    device arrays over the batch (one simulation here)."""

    def __init__(self, n, reorder):
        super().__init__(reorder)
        self.N, self.B, self.dev = n, 1, torch.device("cpu")
        self.cells = np.array([float(n)])

    def tile_order(self, x, order):
        order.copy_(MockStepper.tile_order(self, x))

    def field_kick(self, x, v, gamma, field, chain=False):
        field[0] = x.clone()
        if v is None:
            self.field_only += 1
        else:
            v += gamma * force(x)

    def field_kick_drift(self, x, v, gamma, beta, field, xout, xkeep=None, chain=False, vout=None):
        if vout is not None:
            assert vout is not v
            vout.copy_(v)
            v = vout
        self.field_kick(x, v, gamma, field)
        xout.copy_(x + beta * v)

    def vjp_field_drift(self, field, x, cot, scale, pot_scale, beta, gpos, chain=False):
        self._vjp(field, x, cot, scale, gpos)
        cot += beta * gpos

    def vjp_field(self, field, x, cot, scale, pot_scale, gpos):
        self._vjp(field, x, cot, scale, gpos)


def torch_permute_rows(order, *arrays, scatter=False, out=None):
    idx = order.long()
    outs = list(out) if out is not None else [torch.empty_like(a) for a in arrays]
    for a, o in zip(arrays, outs):
        if scatter:
            o[idx] = a
        else:
            o.copy_(a[idx])
    return tuple(outs) if len(outs) > 1 else outs[0]


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("S,every", [(1, None), (5, 2), (6, 3), (4, 1)])
def test_batch_leapfrog_buffers_and_adjoint(monkeypatch, reorder, S, every):
    monkeypatch.setattr(integrate._ops, "permute_rows", torch_permute_rows)
    k0, beta, gamma = schedule(S, 20 + S)
    r = np.random.default_rng(100 + S)
    n = 36
    x0 = torch.as_tensor(r.uniform(-2, 2, (n, 3)).astype(np.float32))
    v0 = torch.as_tensor(r.normal(size=(n, 3)).astype(np.float32))
    gx = torch.as_tensor(r.normal(size=(n, 3)).astype(np.float32))
    gv = torch.as_tensor(r.normal(size=(n, 3)).astype(np.float32))
    c = np.concatenate([[k0], beta, gamma])[None, :]
    f32 = c.astype(np.float32)[0].astype(np.float64)   # the float32 coefficients the batch applies
    st = MockBatchStepper(n, reorder)
    eng = integrate.BatchLeapfrog(st, c, every)
    ra = x0.double().requires_grad_(True)
    rb = v0.double().requires_grad_(True)
    xr, vr = plain_leapfrog(ra, rb, f32[0], f32[1:1 + S], f32[1 + S:])
    ((xr * gx.double()).sum() + (vr * gv.double()).sum()).backward()
    for rep in range(2):
        a, b = x0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
        x, v = integrate.BatchNBody.apply(a, b, torch.as_tensor(c), eng, every)
        ((x * gx).sum() + (v * gv).sum()).backward()
        torch.testing.assert_close(x.detach().double(), xr.detach(), rtol=1e-5, atol=1e-5)
        torch.testing.assert_close(v.detach().double(), vr.detach(), rtol=1e-5, atol=1e-5)
        torch.testing.assert_close(a.grad.double(), ra.grad, rtol=1e-4, atol=1e-4)
        torch.testing.assert_close(b.grad.double(), rb.grad, rtol=1e-4, atol=1e-4)
    assert st.vjps == 2 * (S + 1)
    # static outputs stay valid through the adjoint, and a second backward
    # without a forward is refused
    xs, vs = eng.forward(x0, v0)
    keep = xs.clone()
    eng.backward(gx, gv)
    assert torch.equal(xs, keep)
    with pytest.raises(RuntimeError):
        eng.backward(gx, gv)
