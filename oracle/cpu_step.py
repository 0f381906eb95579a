"""CPU timing port of one forward + adjoint leapfrog step (test infra only).

The same algorithm as the oracle and the GPU path -- c2c FFTs (the
reference's ``fft3d``/``ifft3d``), scatter-add painting and gather reading
with the corner maths of painting.py:15-106 -- in float32 with
``scipy.fft`` using all host cores.  ``bench.py`` times it as the CPU
baseline; it is checked against :mod:`oracle.adjoint` in the tests.
"""
import numpy as np
import scipy.fft as sfft

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]])


def _corners(pos, shape):
    fl = np.floor(pos)
    fr = (pos - fl).astype(np.float32)
    base = fl.astype(np.int64)
    n = np.array(shape)
    out = []
    for o in CORNERS:
        idx = np.mod(base + o, n)
        flat = (idx[:, 0] * shape[1] + idx[:, 1]) * shape[2] + idx[:, 2]
        k = np.where(o == 1, fr, 1 - fr)
        dk = np.where(o == 1, np.float32(1), np.float32(-1))
        out.append((flat, k, dk))
    return out


def paint(cs, m, w=None):
    out = np.zeros(m, np.float64)
    for flat, k, _ in cs:
        wt = k[:, 0] * k[:, 1] * k[:, 2]
        out += np.bincount(flat, weights=wt if w is None else wt * w, minlength=m)
    return out.astype(np.float32)


def read(cs, mesh):
    flatm = mesh.reshape(-1)
    s = 0
    for flat, k, _ in cs:
        s = s + flatm[flat] * (k[:, 0] * k[:, 1] * k[:, 2])
    return s


def read_grad(cs, mesh):
    flatm = mesh.reshape(-1)
    g = np.zeros((cs[0][1].shape[0], 3), np.float32)
    for flat, k, dk in cs:
        v = flatm[flat]
        g[:, 0] += v * dk[0] * k[:, 1] * k[:, 2]
        g[:, 1] += v * k[:, 0] * dk[1] * k[:, 2]
        g[:, 2] += v * k[:, 0] * k[:, 1] * dk[2]
    return g


def multipliers(shape):
    kv = [(2 * np.pi * np.fft.fftfreq(n)).astype(np.float32).reshape([-1 if i == d else 1 for i in range(3)])
          for d, n in enumerate(shape)]
    kk = kv[0] ** 2 + kv[1] ** 2 + kv[2] ** 2
    green = -np.where(kk == 0, 0, 1 / np.where(kk == 0, 1, kk)).astype(np.float32)
    grads = [(1j * (8 * np.sin(k) - np.sin(2 * k)) / 6).astype(np.complex64) for k in kv]
    return green, grads


def force_field(pos, shape, workers):
    m = int(np.prod(shape))
    cs = _corners(pos, shape)
    rho = paint(cs, m).reshape(shape)
    dk = sfft.fftn(rho, workers=workers)
    green, grads = multipliers(shape)
    pot = dk * green
    phis = [sfft.ifftn(-g * pot, workers=workers).real.astype(np.float32) for g in grads]
    return cs, phis


def step_fwd_adj(x, v, cot_x, cot_v, shape, beta, gamma, workers):
    """Drift + force + kick, a recomputed force (checkpointed adjoint), and
    the adjoint of the kick and drift.  Returns the new (x, v, lx, lv)."""
    m = int(np.prod(shape))
    x = x + beta * v
    cs, phis = force_field(x, shape, workers)
    v = v + gamma * np.stack([read(cs, p) for p in phis], -1)
    cs, phis = force_field(x, shape, workers)  # recompute in the backward sweep
    green, grads = multipliers(shape)
    u = gamma * cot_v
    g = sum(u[:, d:d + 1] * read_grad(cs, phis[d]) for d in range(3))
    rk = 0
    for d in range(3):
        rk = rk + np.conj(-grads[d] * green) * sfft.fftn(paint(cs, m, u[:, d]).reshape(shape), workers=workers)
    rho_bar = sfft.ifftn(rk, workers=workers).real.astype(np.float32)
    lx = cot_x + g + read_grad(cs, rho_bar)
    lv = cot_v + beta * lx
    return x, v, lx, lv
