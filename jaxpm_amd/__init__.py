"""jaxpm_amd: the JaxPM particle-mesh step on MI355X (gfx950).

Modules mirror the reference package layout (``jaxpm.painting``,
``jaxpm.kernels``, ``jaxpm.pm``, ``jaxpm.ode``, ``jaxpm.growth``,
``jaxpm.distributed``, ``jaxpm.utils``); the numerical work runs in
``libjaxpm_hip.so`` (hand-written HIP kernels + hipFFT), loaded lazily on
first use.  Build it with ``python -m jaxpm_amd.build``.
"""
from . import config, cosmology, growth  # noqa: F401 (host-only, import eagerly)

__all__ = ["config", "painting", "kernels", "pm", "ode", "growth", "cosmology", "distributed", "utils", "integrate"]


def __getattr__(name):
    if name in __all__:
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
