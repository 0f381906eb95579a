"""Input normalisation shared by the public modules."""
import numbers

import numpy as np
import torch

from . import _lib


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
