"""The allocation-free binned paint (jpm_cic_paint_binned_noalloc_f32), called
through the C ABI, against jpm_cic_paint_binned_f32 with a caller workspace:
deterministic mode gives the same bits, on two calls in a row, the second on
a smaller input (the scratch keeps the first call's contents), plus the 3-channel
form, the oracle, and the host-side rejection of inputs beyond the scratch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import painting as op

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20   # JPM_NOALLOC_SCRATCH_BYTES


def _paint_pair(lib_mod, shape, pos, w, nc, flags):
    """(with workspace, without) meshes for one input."""
    from jaxpm_amd._lib import call, ptr, size, stream
    nx, ny, nz = shape
    n = pos.shape[0]
    m = nx * ny * nz
    ref = torch.full((nc * m,), 7.0, device="cuda")     # ZERO_MESH overwrites it
    got = torch.full((nc * m,), -3.0, device="cuda")
    wsb = size("jpm_cic_paint_binned_workspace_size", nx, ny, nz, n, nc, flags)
    assert 0 < wsb <= SCRATCH
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    call("jpm_cic_paint_binned_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(w), 1.0, nc, ptr(ref), flags, ptr(ws), wsb)
    call("jpm_cic_paint_binned_noalloc_f32", stream(), nx, ny, nz, ptr(pos), n, ptr(w), 1.0, nc, ptr(got), flags)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_paint_matches_workspace_paint(gpu):
    from jaxpm_amd import _lib
    flags = _lib.DETERMINISTIC | _lib.ZERO_MESH
    r = np.random.default_rng(71)
    # first a 64^3 mesh with 64^3 clustered particles and weights, then a
    # smaller ragged mesh: the second call runs on the scratch the first left
    cases = [((64, 64, 64), 64 ** 3, True), ((20, 24, 36), 9000, False)]
    for shape, n, clustered in cases:
        print(f"noalloc paint {shape} n={n}", flush=True)
        L = np.array(shape, dtype=np.float64)
        if clustered:
            pos = (L / 2 + r.normal(scale=4.0, size=(n, 3))) % L
        else:
            pos = r.uniform(-0.5, 1.5, size=(n, 3)) * L
        pos = pos.astype(np.float32)
        w = r.uniform(0.5, 2.0, size=n).astype(np.float32)
        pt = torch.as_tensor(pos, device="cuda")
        wt = torch.as_tensor(w, device="cuda")
        ref, got = _paint_pair(_lib, shape, pt, wt, 1, flags)
        assert np.array_equal(ref, got), f"{shape}: noalloc differs from the workspace paint"
        want = op.cic_paint(np.zeros(shape), pos, w)
        err = np.abs(got.reshape(shape) - want).max() / np.abs(want).max()
        assert err < 2e-6, f"{shape}: error {err:.3e} against the oracle"
        # scalar weight (weights NULL) on the same scratch
        ref, got = _paint_pair(_lib, shape, pt, None, 1, flags)
        assert np.array_equal(ref, got)


def test_noalloc_paint_three_channels(gpu):
    from jaxpm_amd import _lib
    flags = _lib.DETERMINISTIC | _lib.ZERO_MESH
    r = np.random.default_rng(72)
    for shape, n in [((32, 32, 32), 40000), ((16, 16, 16), 3000)]:
        print(f"noalloc paint nc=3 {shape} n={n}", flush=True)
        pos = (r.uniform(0, 1, size=(n, 3)) * np.array(shape)).astype(np.float32)
        w = r.normal(size=(n, 3)).astype(np.float32)
        ref, got = _paint_pair(_lib, shape, torch.as_tensor(pos, device="cuda"), torch.as_tensor(w, device="cuda"),
                               3, flags)
        assert np.array_equal(ref, got)
        for c in range(3):
            want = op.cic_paint(np.zeros(shape), pos, w[:, c])
            err = np.abs(got.reshape(3, *shape)[c] - want).max() / np.abs(want).max()
            assert err < 2e-6


def test_noalloc_paint_rejects_inputs_beyond_scratch(gpu):
    """Checked on the host before any launch: nothing runs on the device."""
    from jaxpm_amd._lib import load, size
    lib = load()
    nx = ny = nz = 256
    n = 256 ** 3
    assert size("jpm_cic_paint_binned_workspace_size", nx, ny, nz, n, 1, 2) > SCRATCH
    dummy = torch.zeros(16, device="cuda")
    rc = lib.jpm_cic_paint_binned_noalloc_f32(None, nx, ny, nz, ctypes.c_void_p(dummy.data_ptr()), n, None, 1.0, 1,
                                              ctypes.c_void_p(dummy.data_ptr()), 2)
    assert rc == -1
    assert b"scratch" in lib.jpm_last_error()
    rc = lib.jpm_cic_paint_binned_noalloc_f32(None, 8, 8, 8, ctypes.c_void_p(dummy.data_ptr()), 4, None, 1.0, 2,
                                              ctypes.c_void_p(dummy.data_ptr()), 2)
    assert rc == -1 and b"channels" in lib.jpm_last_error()
