"""Process-wide switches, mirroring ``jax.config`` for the one the
reference's tests use: ``update("jax_enable_x64", True)``
(tests/test_distributed_pm.py:30), also read from the ``JAX_ENABLE_X64``
environment variable at import, as JAX does.

With it off (the default, as in JAX), float64 input is computed in float32.
With it on, float64 input to ``cic_paint``, ``cic_read``, ``cic_paint_dx``,
``cic_read_dx`` and ``pm_forces`` runs the float64 kernels
(``jpm_cic_paint_f64``, ``jpm_cic_read_f64``, ``jpm_pm_forces_f64``) and
returns float64; those paths are forward-only.
"""
import os

__all__ = ["update", "x64_enabled"]

_flags = {"jax_enable_x64": os.environ.get("JAX_ENABLE_X64", "0").lower() in ("1", "true", "yes", "on")}


def update(name, value):
    if name not in _flags:
        raise ValueError(f"unknown option {name!r}; known: {sorted(_flags)}")
    _flags[name] = bool(value)


def x64_enabled():
    return _flags["jax_enable_x64"]
