"""Pins for oracle.painting: brute force, identities, transposes."""
import numpy as np
import pytest

from oracle import painting as op


def brute_paint(shape, positions, weight=None):
    """Per particle, per corner loop in plain Python (the slow pin)."""
    out = np.zeros(shape)
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    for p in range(pos.shape[0]):
        wp = 1.0 if weight is None else float(np.ravel(weight)[p])
        base = [int(np.floor(x)) for x in pos[p]]
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    c = (base[0] + dx, base[1] + dy, base[2] + dz)
                    w = 1.0
                    for d in range(3):
                        w *= 1.0 - abs(pos[p, d] - c[d])
                    out[c[0] % shape[0], c[1] % shape[1], c[2] % shape[2]] += wp * w
    return out


def brute_read(mesh, positions):
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    shape = mesh.shape
    out = np.zeros(pos.shape[0])
    for p in range(pos.shape[0]):
        base = [int(np.floor(x)) for x in pos[p]]
        s = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    c = (base[0] + dx, base[1] + dy, base[2] + dz)
                    w = 1.0
                    for d in range(3):
                        w *= 1.0 - abs(pos[p, d] - c[d])
                    s += w * mesh[c[0] % shape[0], c[1] % shape[1], c[2] % shape[2]]
        out[p] = s
    return out


@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 6, 5)])
def test_paint_matches_brute_force_with_wrap(shape):
    r = np.random.default_rng(0)
    # include negative positions and positions beyond the box
    pos = r.uniform(-1.5 * max(shape), 2.5 * max(shape), size=(60, 3))
    w = r.normal(size=60)
    got = op.cic_paint(np.zeros(shape), pos, w)
    np.testing.assert_allclose(got, brute_paint(shape, pos, w), rtol=0, atol=1e-12)


def test_read_matches_brute_force():
    r = np.random.default_rng(1)
    mesh = r.normal(size=(8, 6, 5))
    pos = r.uniform(-10, 20, size=(50, 3))
    np.testing.assert_allclose(op.cic_read(mesh, pos), brute_read(mesh, pos), rtol=0, atol=1e-12)


def test_paint_accumulates_into_given_mesh():
    r = np.random.default_rng(2)
    base = r.normal(size=(4, 4, 4))
    pos = r.uniform(0, 4, size=(10, 3))
    np.testing.assert_allclose(op.cic_paint(base, pos) - base, op.cic_paint(np.zeros((4, 4, 4)), pos),
                               atol=1e-12)


def test_mass_conservation_and_linearity():
    r = np.random.default_rng(3)
    pos = r.uniform(-8, 24, size=(500, 3))
    w1, w2 = r.normal(size=500), r.normal(size=500)
    m1 = op.cic_paint(np.zeros((16, 16, 16)), pos, w1)
    m2 = op.cic_paint(np.zeros((16, 16, 16)), pos, w2)
    assert abs(m1.sum() - w1.sum()) < 1e-10
    np.testing.assert_allclose(op.cic_paint(np.zeros((16, 16, 16)), pos, 2 * w1 - 3 * w2), 2 * m1 - 3 * m2,
                               atol=1e-10)


def test_lattice_paints_one_per_cell_and_reads_node_values():
    shape = (6, 5, 4)
    lattice = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing='ij'), -1).reshape(-1, 3)
    np.testing.assert_array_equal(op.cic_paint(np.zeros(shape), lattice), np.ones(shape))
    mesh = np.random.default_rng(4).normal(size=shape)
    np.testing.assert_array_equal(op.cic_read(mesh, lattice), mesh.reshape(-1))
    np.testing.assert_array_equal(op.cic_paint_dx(np.zeros((*shape, 3))), np.ones(shape))
    np.testing.assert_array_equal(op.cic_read_dx(mesh, np.zeros((*shape, 3))), mesh)


def test_integer_shift_equals_roll():
    r = np.random.default_rng(5)
    shape = (8, 8, 8)
    pos = r.uniform(0, 8, size=(100, 3))
    m = op.cic_paint(np.zeros(shape), pos)
    shifted = op.cic_paint(np.zeros(shape), pos + np.array([3, -2, 9]))
    np.testing.assert_allclose(shifted, np.roll(m, (3, -2, 9), axis=(0, 1, 2)), atol=1e-12)


def test_paint_read_transpose_identity():
    """<paint(x, w), g> == <w, read(g, x)> exactly in float64."""
    r = np.random.default_rng(6)
    shape = (8, 7, 6)
    pos = r.uniform(-5, 15, size=(200, 3))
    w = r.normal(size=200)
    g = r.normal(size=shape)
    lhs = np.sum(op.cic_paint(np.zeros(shape), pos, w) * g)
    rhs = np.sum(w * op.cic_read(g, pos))
    assert abs(lhs - rhs) < 1e-10 * (abs(lhs) + 1)


def test_absolute_equals_relative():
    r = np.random.default_rng(7)
    shape = (8, 8, 8)
    disp = r.normal(scale=2.0, size=(*shape, 3))
    lattice = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing='ij'), -1)
    np.testing.assert_allclose(op.cic_paint_dx(disp), op.cic_paint(np.zeros(shape), lattice + disp), atol=1e-10)
    mesh = r.normal(size=shape)
    np.testing.assert_allclose(op.cic_read_dx(mesh, disp), op.cic_read(mesh, lattice + disp).reshape(shape),
                               atol=1e-10)


def test_relative_float32_drops_corner_that_rounds_onto_L():
    """A displacement of -1e-9 at node 0 wraps to L in float32 (jnp.mod rounds
    to L), the corner index L is out of range and dropped by the reference's
    scatter -- so a little mass is lost exactly as in the reference."""
    shape = (4, 4, 4)
    disp = np.zeros((*shape, 3), np.float32)
    disp[0, 0, 0] = (-1e-9, 0, 0)
    idx, w, valid = op.rel_corners(disp, np.float32)
    assert not valid[:, 0].all()
    m = op.cic_paint_dx(disp, dtype=np.float32)
    assert m.sum() < 64.0


def test_compensate_cic_plane_wave():
    n = 16
    x = np.arange(n)
    f = np.cos(2 * np.pi * 3 * x / n)[:, None, None] * np.ones((n, n, n))
    k = 2 * np.pi * 3 / n
    expect = f / np.sinc(k / (2 * np.pi)) ** 2
    np.testing.assert_allclose(op.compensate_cic(f), expect, atol=1e-10)


def test_weight_shape_error():
    with pytest.raises(ValueError, match="Weight shape must match particle shape"):
        op.cic_paint_dx(np.zeros((4, 4, 4, 3)), weight=np.ones(5))
