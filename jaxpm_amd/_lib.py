"""ctypes binding of libjaxpm_hip.so (C ABI declared in include/jaxpm_hip.h).

Loading fails loudly when the library is not built; there is no CPU
fallback anywhere in the package.
"""
import ctypes
import os
import threading

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libjaxpm_hip.so")

# Flags, mirrored from include/jaxpm_hip.h.
RELATIVE = 1
DETERMINISTIC = 2
ZERO_MESH = 4
ATOMIC = 8
K_GRAD_EXACT = 1
K_LAPL_FD = 2
K_DECONV = 4


def tune_flags(chunk_code=0, slice_code=0, rho_global=0, xcd_slabs=0, faces_in_mesh=0):
    """Sort / gather tuning bits of the pipeline flags (JPM_TUNE_CHUNK,
    JPM_TUNE_SLICE, JPM_TUNE_RHO_GLOBAL, JPM_TUNE_XCD_SLABS,
    JPM_TUNE_FACES_IN_MESH): chunk = 2^(9 + chunk_code), slice = 2^(11 +
    slice_code), 0 = default; rho_global reads the force-VJP density adjoint
    through the caches instead of LDS; xcd_slabs gives each XCD one
    contiguous run of tiles; faces_in_mesh has the force evaluations' paint
    write its tiles' face cells into the mesh instead of a buffer of their
    own (the same bits)."""
    return ((int(chunk_code) << 16) | (int(slice_code) << 20) | (int(bool(rho_global)) << 24)
            | (int(bool(xcd_slabs)) << 25) | (int(bool(faces_in_mesh)) << 26))

E_INVALID = -1

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_f = ctypes.c_float
_sz = ctypes.c_size_t
_d = ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "jpm_init": (_i, []),
    "jpm_shutdown": (_i, []),
    "jpm_abi_version": (_i, []),
    "jpm_last_error": (ctypes.c_char_p, []),
    "jpm_timing_enable": (_i, [_i]),
    "jpm_timing_report": (_i, [ctypes.c_char_p, _sz]),
    "jpm_cic_paint_workspace_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_cic_paint_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _f, _vp, _u32, _vp, _sz]),
    "jpm_cic_read_f32": (_i, [_vp, _i, _i, _i, _vp, _i, _i64, _vp, _i64, _vp, _u32]),
    "jpm_cic_read_sorted_workspace_size": (_sz, [_i, _i, _i, _i64]),
    "jpm_cic_read_sorted_f32": (_i, [_vp, _i, _i, _i, _vp, _i, _i64, _vp, _i64, _vp, _u32, _vp, _sz]),
    "jpm_cic_paint_multi_workspace_size": (_sz, [_i, _i, _i, _i, _u32]),
    "jpm_cic_paint_multi_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _i, _vp, _u32, _vp, _sz]),
    "jpm_cic_grad_pos_f32": (_i, [_vp, _i, _i, _i, _vp, _i, _i64, _vp, _i64, _vp, _vp, _f, _vp, _u32, _i]),
    "jpm_cic_paint_f64": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _d, _vp, _u32]),
    "jpm_cic_read_f64": (_i, [_vp, _i, _i, _i, _vp, _i, _i64, _vp, _i64, _vp, _u32]),
    "jpm_pm_forces_f64_workspace_size": (_sz, [_i, _i, _i]),
    "jpm_pm_forces_f64": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _d, _u32, _u32, _vp, _vp, _sz]),
    "jpm_fft_r2c_f32": (_i, [_vp, _i, _i, _i, _i, _vp, _vp]),
    "jpm_fft_c2r_f32": (_i, [_vp, _i, _i, _i, _i, _vp, _vp]),
    "jpm_fft_c2c_workspace_size": (_sz, [_i, _i, _i, _i]),
    "jpm_fft_c2c_f32": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i, _vp, _sz]),
    "jpm_kspace_force_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _f, _u32]),
    "jpm_kspace_force_adjoint_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _f, _u32]),
    "jpm_kspace_compensate_f32": (_i, [_vp, _i, _i, _i, _vp, _f]),
    "jpm_kspace_potential_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _f, _u32]),
    "jpm_fd4_gradient_f32": (_i, [_vp, _i, _i, _i, _vp, _f, _vp]),
    "jpm_fd4_divergence_f32": (_i, [_vp, _i, _i, _i, _vp, _f, _vp]),
    "jpm_kspace_linear_field_f32": (_i, [_vp, _i, _i, _i, _vp, _f, _f, _f, _vp, _vp, _i, _f]),
    "jpm_kspace_shear6_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _u32]),
    "jpm_lpt2_source_f32": (_i, [_vp, _i64, _vp, _vp]),
    "jpm_kspace_shear6_adjoint_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _u32]),
    "jpm_lpt2_source_vjp_f32": (_i, [_vp, _i64, _vp, _vp, _vp]),
    "jpm_spectrum_full_to_half_f32": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "jpm_spectrum_half_to_full_f32": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "jpm_linear_field_table_vjp_workspace_size": (_sz, [_i]),
    "jpm_kspace_linear_field_table_vjp_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _f, _f, _vp, _vp, _i, _f, _vp, _vp,
                                                   _sz]),
    "jpm_dot_workspace_size": (_sz, []),
    "jpm_dot_f32": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _sz]),
    "jpm_batch_create": (_i, [_i, _vp, _vp, _u32, ctypes.POINTER(_vp)]),
    "jpm_batch_destroy": (_i, [_vp]),
    "jpm_batch_sort_size": (_sz, [_vp]),
    "jpm_batch_force_kick_workspace_size": (_sz, [_vp]),
    "jpm_batch_force_vjp_workspace_size": (_sz, [_vp]),
    "jpm_batch_dot_workspace_size": (_sz, [_vp]),
    "jpm_batch_force_kick_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _u32, _i, _vp, _sz, _vp, _sz, _vp]),
    "jpm_batch_force_vjp_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _u32, _vp, _sz, _vp, _sz]),
    "jpm_batch_force_kick_drift_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _u32, _vp, _sz, _vp,
                                            _sz, _vp, _vp, _vp]),
    "jpm_batch_tile_count": (_sz, [_vp]),
    "jpm_batch_force_vjp_drift_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _u32, _vp, _sz, _vp,
                                           _sz, _vp, _vp]),
    "jpm_batch_axpy_f32": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "jpm_batch_tile_order_u32": (_i, [_vp, _vp, _vp, _vp, _vp, _sz]),
    "jpm_batch_dot_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _sz]),
    "jpm_batch_lpt_workspace_size": (_sz, [_vp]),
    "jpm_batch_linear_field_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _sz]),
    "jpm_batch_lattice_forces_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _sz]),
    "jpm_batch_lattice_forces_vjp_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _sz]),
    "jpm_batch_paint_f32": (_i, [_vp, _vp, _vp, _vp, _vp, _sz]),
    "jpm_batch_paint_vjp_f32": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "jpm_power_spectrum_workspace_size": (_sz, [_i, _i, _i, _i, _i, _i]),
    "jpm_power_spectrum_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _d, _d, _d, _vp, _i, _vp, _i, _d, _d, _d, _vp, _vp,
                                    _sz]),
    "jpm_force_field_workspace_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_force_field_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _vp, _f, _f, _u32, _vp, _sz]),
    "jpm_force_vjp_workspace_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_force_vjp_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _vp, _f, _vp, _f, _u32, _vp, _sz]),
    "jpm_force_kick_workspace_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_force_kick_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _f, _vp, _f, _u32, _i, _vp, _sz, _vp, _sz, _vp]),
    "jpm_bin_sort_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_force_vjp_pot_workspace_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_force_vjp_pot_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _vp, _f, _vp, _f, _u32, _vp, _sz, _vp, _sz]),
    "jpm_force_kick_drift_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _f, _f, _vp, _vp, _vp, _f, _u32, _vp, _sz,
                                      _vp, _sz, _vp, _vp, _vp]),
    "jpm_tile_count": (_sz, [_i, _i, _i]),
    "jpm_force_vjp_pot_drift_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _vp, _f, _f, _vp, _f, _u32, _vp, _sz, _vp,
                                         _sz, _vp, _vp]),
    "jpm_axpy_f32": (_i, [_vp, _i64, _f, _vp, _vp]),
    "jpm_cic_paint_2d_f32": (_i, [_vp, _i, _i, _vp, _i64, _i, _vp, _vp]),
    "jpm_cic_paint_2d_vjp_f32": (_i, [_vp, _i, _i, _vp, _i64, _i, _vp, _vp, _vp, _i, _vp]),
    "jpm_density_plane_f32": (_i, [_vp, _vp, _i64, _i, _i, _d, _d, _i, _vp]),
    "jpm_density_plane_vjp_f32": (_i, [_vp, _vp, _i64, _i, _i, _d, _d, _i, _vp, _vp]),
    "jpm_gaussian_smoothing_workspace_size": (_sz, [_i, _i]),
    "jpm_gaussian_smoothing_f32": (_i, [_vp, _i, _i, _vp, _vp, _d, _vp, _sz]),
    "jpm_tile_order_workspace_size": (_sz, [_i, _i, _i, _i64]),
    "jpm_tile_order_u32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _vp, _sz]),
    "jpm_permute_rows_f32": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _i]),
    "jpm_kick_f32": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _f, _vp, _u32]),
    "jpm_pm_forces_workspace_size": (_sz, [_i, _i, _i, _i64, _u32]),
    "jpm_cic_paint_binned_workspace_size": (_sz, [_i, _i, _i, _i64, _i, _u32]),
    "jpm_cic_paint_binned_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _f, _i, _vp, _u32, _vp, _sz]),
    "jpm_cic_paint_binned_noalloc_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _f, _i, _vp, _u32]),
    "jpm_pm_forces_f32": (_i, [_vp, _i, _i, _i, _vp, _i64, _vp, _vp, _f, _f, _u32, _vp, _sz]),
}

_lock = threading.Lock()
_lib = None


class JpmError(RuntimeError):
    """A libjaxpm_hip call failed on the device side."""


def load():
    """Load the shared library (no device access); raises if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `python -m jaxpm_amd.build` "
                    "(jaxpm_amd has no CPU fallback)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def call(name, *args):
    """Call an entry point and raise ValueError / JpmError on failure."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        msg = load().jpm_last_error().decode(errors="replace")
        if rc == E_INVALID:
            raise ValueError(f"{name}: {msg}")
        raise JpmError(f"{name} failed ({rc}): {msg}")
    return rc


def size(name, *args):
    return int(getattr(load(), name)(*args))


_device_ready = False


def device():
    """The HIP device the library runs on; raises if none is available."""
    global _device_ready
    if not torch.cuda.is_available():
        raise RuntimeError("jaxpm_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    if not _device_ready:
        call("jpm_init")
        _device_ready = True
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    """Raw hipStream_t of torch's current stream."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


_timing_on = False


def timing(enable):
    """Turn the library's per-kernel HIP-event timing on or off."""
    global _timing_on
    call("jpm_timing_enable", int(bool(enable)))
    _timing_on = bool(enable)


def timing_enabled():
    return _timing_on


def timing_report():
    """{kernel: (launches, total_ms)} since the last report (waits for the
    recorded events)."""
    buf = ctypes.create_string_buffer(1 << 16)
    call("jpm_timing_report", buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.rsplit(" ", 2)
        out[name] = (int(cnt), float(ms))
    return out
