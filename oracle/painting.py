"""Oracle CIC painting and reading (test infrastructure only).

Restates JaxPM ``jaxpm/painting.py`` and ``jaxpm/painting_utils.py`` on the
CPU with NumPy.  Accumulation uses ``np.bincount`` (float64 sums), so the
oracle is order-independent up to float64 rounding.
"""
import numpy as np

# Corner order of the absolute painter, painting.py:24-25.
ABS_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                        [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]])
# Corner order of the relative painter: bit d of the corner number is the
# offset along axis d, painting_utils.py:43-45.
REL_CORNERS = np.array([[(c >> d) & 1 for d in range(3)] for c in range(8)])


def _flat(idx, shape):
    return (idx[:, 0] * shape[1] + idx[:, 1]) * shape[2] + idx[:, 2]


def abs_corners(positions, shape, dtype=np.float64):
    """Per-corner flat indices and weights of absolute CIC.

    painting.py:21-37: floor, corner = floor + offset, weight
    prod(1 - |pos - corner|), int32 cast of the corner, then ``mod shape``.
    Returns (idx [8, P] int64, w [8, P] dtype).  The per-axis factors are
    formed once for each offset and multiplied in axis order, the same float
    operations per corner as forming them corner by corner.
    """
    pos = np.asarray(positions, dtype).reshape(-1, 3)
    fl = np.floor(pos)
    kk, cc = [], []
    for d in range(3):
        kd, cd = [], []
        for o in (0, 1):
            c = fl[:, d] + dtype(o)
            kd.append(dtype(1) - np.abs(pos[:, d] - c))
            cd.append(np.mod(c.astype(np.int32).astype(np.int64), np.int64(shape[d])))
        kk.append(kd)
        cc.append(cd)
    idxs, ws = [], []
    for o in ABS_CORNERS:
        ws.append(kk[0][o[0]] * kk[1][o[1]] * kk[2][o[2]])
        idxs.append((cc[0][o[0]] * shape[1] + cc[1][o[1]]) * shape[2] + cc[2][o[2]])
    return np.stack(idxs), np.stack(ws)


def rel_corners(displacements, dtype=np.float64):
    """Per-corner indices and weights of relative CIC (``enmesh``).

    painting_utils.py:43-69 with ``new_cell_size`` given: position
    ``p = lattice + disp`` (48), ``q = (p + o) mod L`` (51-55), ``floor`` (57),
    ``nd = p - q`` rewrapped with ``rint(nd / L) * L`` (58-63), weight
    ``prod(1 - |nd|)`` (64, 69).  Corners whose index equals L after float
    rounding are out of range; the reference's scatter drops them and its
    gather (mode='drop') reads 0, so they are returned with ``valid = False``.
    Returns (idx [8, P] int64, w [8, P], valid [8, P] bool).
    """
    disp = np.asarray(displacements, dtype)
    shape = disp.shape[:-1]
    grid = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing='ij'),
                    axis=-1).reshape(-1, 3)
    p = grid.astype(dtype) + disp.reshape(-1, 3)
    L = np.array(shape, dtype)
    shp = np.array(shape, np.int64)
    idxs, ws, vs = [], [], []
    for o in REL_CORNERS:
        q = np.mod(p + o.astype(dtype), L)
        q = np.floor(q)
        nd = p - q
        nd = nd - np.rint(nd / L) * L
        k = dtype(1) - np.abs(nd)
        ws.append(k[:, 0] * k[:, 1] * k[:, 2])
        qi = q.astype(np.int64)
        valid = np.all((qi >= 0) & (qi < shp), axis=-1)
        qi = np.where(valid[:, None], qi, 0)
        idxs.append(_flat(qi, shape))
        vs.append(valid)
    return np.stack(idxs), np.stack(ws), np.stack(vs)


def _deposit(shape, idx, w):
    m = int(np.prod(shape))
    out = np.zeros(m, np.float64)
    for c in range(idx.shape[0]):
        out += np.bincount(idx[c], weights=w[c].astype(np.float64), minlength=m)
    return out.reshape(shape)


def cic_paint(grid_mesh, positions, weight=1.0, dtype=np.float64):
    """Absolute CIC paint, painting.py:15-45 (``_cic_paint_impl``).

    Accumulates into a copy of ``grid_mesh`` (scatter_add into the given
    mesh, painting.py:43-44).  ``weight`` is a scalar or one value per
    particle (painting.py:30-33).
    """
    mesh = np.asarray(grid_mesh)
    shape = mesh.shape
    idx, w = abs_corners(positions, shape, dtype)
    wt = np.asarray(weight, dtype).reshape(-1)
    w = w * (wt if wt.size > 1 else wt[0])
    return (mesh.astype(np.float64) + _deposit(shape, idx, w)).astype(dtype)


def cic_read(grid_mesh, positions, dtype=np.float64, corners=None):
    """Absolute CIC read, painting.py:78-106: sum_o mesh[idx_o] * W_o.
    ``corners``: the (idx, w) of ``abs_corners`` for these positions, when
    several meshes are read at the same positions."""
    mesh = np.asarray(grid_mesh, dtype)
    pos = np.asarray(positions)
    idx, w = abs_corners(pos, mesh.shape, dtype) if corners is None else corners
    vals = mesh.reshape(-1)[idx].astype(np.float64) * w
    return vals.sum(0).astype(dtype).reshape(pos.shape[:-1])


def cic_paint_dx(displacements, weight=1.0, dtype=np.float64):
    """Relative CIC paint, painting.py:161-189 + painting_utils.scatter.

    Particles sit at lattice node (i, j, k) plus ``displacements[i, j, k]``;
    the mesh has the lattice's shape and starts at zero (painting.py:170).
    """
    disp = np.asarray(displacements)
    shape = disp.shape[:-1]
    idx, w, valid = rel_corners(disp, dtype)
    wt = np.asarray(weight, dtype).reshape(-1)
    if wt.size > 1 and wt.size != int(np.prod(shape)):
        raise ValueError("Weight shape must match particle shape")
    # dropped corners deposit nothing, not even a NaN weight (the
    # scatter's out-of-bounds mode drops the update)
    w = np.where(valid, w * (wt if wt.size > 1 else wt[0]), 0)
    return _deposit(shape, idx, w).astype(dtype)


def cic_read_dx(grid_mesh, disp, dtype=np.float64):
    """Relative CIC read, painting.py:218-236 + painting_utils.gather."""
    mesh = np.asarray(grid_mesh, dtype)
    idx, w, valid = rel_corners(np.asarray(disp).reshape(*mesh.shape, 3), dtype)
    vals = mesh.reshape(-1)[idx].astype(np.float64) * w * valid
    return vals.sum(0).astype(dtype).reshape(mesh.shape)




# Corner order of the 2-D painter, painting.py:141.
CORNERS_2D = np.array([[0, 0], [1, 0], [0, 1], [1, 1]])


def cic_paint_2d(mesh, positions, weight, dtype=np.float64):
    """Weighted 2-D CIC paint, painting.py:131-158: corners floor + offset,
    kernel (1 - |x - cx|)(1 - |y - cy|) times ``weight`` (when not None),
    int32 cast then ``mod mesh.shape``, scatter_add into a copy of ``mesh``."""
    m = np.asarray(mesh)
    pos = np.asarray(positions, dtype).reshape(-1, 2)
    fl = np.floor(pos)
    shp = np.array(m.shape, np.int64)
    out = m.astype(np.float64).reshape(-1).copy()
    wt = None if weight is None else np.asarray(weight, dtype).reshape(-1)
    for o in CORNERS_2D:
        c = fl + o.astype(dtype)
        k = dtype(1) - np.abs(pos - c)
        k = k[:, 0] * k[:, 1]
        if wt is not None:
            k = k * wt
        ci = np.mod(c.astype(np.int32).astype(np.int64), shp)
        out += np.bincount(ci[:, 0] * shp[1] + ci[:, 1], weights=k.astype(np.float64), minlength=out.size)
    return out.reshape(m.shape).astype(dtype)


def compensate_cic(field):
    """painting.py:263-275: ifft(fft(field) * cic_compensation(kvec)).real."""
    from .kernels import cic_compensation, fftk
    f = np.asarray(field, np.float64)
    fk = np.fft.fftn(f)
    return np.fft.ifftn(fk * cic_compensation(fftk(f.shape))).real
