"""The C ABI library loads and exports every entry point include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(jpm_\w+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "jpm_cic_paint_f32" in names and "jpm_pm_forces_f32" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from jaxpm_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libjaxpm_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    # every declared symbol has a Python signature, and vice versa
    assert set(_lib.SIGNATURES) == declared_symbols()
    assert lib.jpm_abi_version() == 1


def test_error_reporting_without_device_work():
    """A bad shape is rejected on the host before any launch (CPU only: the
    calls pass placeholder pointers)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("host-validation test runs without a GPU")
    from jaxpm_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libjaxpm_hip.so not built")
    lib = _lib.load()
    rc = lib.jpm_cic_read_f32(None, 0, 4, 4, None, 1, 0, None, 0, None, 0)
    assert rc == -1
    assert b"mesh shape" in lib.jpm_last_error()
    rc = lib.jpm_cic_read_f32(None, 4, 4, 4, None, 7, 64, None, 0, None, 0)
    assert rc == -1 and b"ncomp" in lib.jpm_last_error()
    # relative mode requires one particle per lattice node
    rc = lib.jpm_cic_paint_f32(None, 4, 4, 4, ctypes.c_void_p(8), 10, None, 1.0, ctypes.c_void_p(8), 1, None, 0)
    assert rc == -1 and b"relative" in lib.jpm_last_error()


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from jaxpm_amd import painting
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        painting.cic_paint(torch.zeros(4, 4, 4), torch.zeros(10, 3))
