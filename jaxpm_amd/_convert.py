"""Input normalisation shared by the public modules."""
import numbers
import warnings

import numpy as np
import torch

from . import _lib, config

_warned_f64 = False


def _is_f64(x):
    if isinstance(x, torch.Tensor):
        return x.dtype == torch.float64
    return isinstance(x, np.ndarray) and x.dtype == np.float64


def wants_f64(*xs):
    """True when x64 is enabled and any of ``xs`` is a float64 array (the
    float64 kernels run).  With x64 off, float64 arrays are computed in
    float32, as JAX does; the first time, a warning says so."""
    global _warned_f64
    if not any(_is_f64(x) for x in xs):
        return False
    if config.x64_enabled():
        return True
    if not _warned_f64:
        _warned_f64 = True
        warnings.warn("float64 input is computed in float32 because jax_enable_x64 is off; "
                      "jaxpm_amd.config.update('jax_enable_x64', True) runs the float64 kernels", stacklevel=3)
    return False


def forward_only(*xs):
    """The float64 paths have no adjoint: refuse inputs that need one."""
    if torch.is_grad_enabled() and any(isinstance(x, torch.Tensor) and x.requires_grad for x in xs):
        raise NotImplementedError("the float64 (jax_enable_x64) paths are forward-only; "
                                  "differentiate in float32")


def to_device(x, dtype=torch.float32):
    """Array-like -> contiguous device tensor of ``dtype`` (keeps autograd)."""
    dev = _lib.device()
    if isinstance(x, torch.Tensor):
        t = x
        if t.device != dev:
            t = t.to(dev)
        if t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()
    arr = np.asarray(x)
    if np.iscomplexobj(arr) and dtype != torch.complex64:
        raise TypeError("complex input where a real array is expected")
    return torch.as_tensor(arr, device=dev).to(dtype).contiguous()


def is_scalar(w):
    if isinstance(w, numbers.Number):
        return True
    if isinstance(w, torch.Tensor):
        return w.dim() == 0
    if isinstance(w, np.ndarray):
        return w.ndim == 0
    return False


def scalar_value(w):
    if isinstance(w, torch.Tensor):
        return float(w.item())
    return float(w)


def single_device(sharding):
    """The reference ignores halo_size without sharding (distributed.py:45-50);
    this build runs on one GPU and has no sharded mode."""
    if sharding is not None:
        raise NotImplementedError(
            "jaxpm_amd runs on one GPU: sharded painting/FFTs (sharding=...) are not supported")
