"""The C-ABI library loads without a GPU and exports every function include/*.h declares."""
import ctypes
import os
import re

from salmon2_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="salmon_hip.h", prefix="salmon_hip_"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", txt)))


def all_declared():
    return declared() + declared("salmon_unet.h", "salmon_unet_") + declared("salmon_pp.h", "salmon_pp_")


def test_headers_are_all_covered():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == [
        "salmon_hip.h", "salmon_pp.h", "salmon_unet.h"]


def test_header_declares_the_binding_list():
    from salmon2_amd import unet

    assert set(declared()) == set(hip.EXPORTS)
    assert set(declared("salmon_unet.h", "salmon_unet_")) == set(unet.EXPORTS)
    from salmon2_amd import pp

    assert set(declared("salmon_pp.h", "salmon_pp_")) == set(pp.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hip.lib_path())
    missing = [n for n in all_declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially():
    assert hip.load().salmon_hip_last_error() in (b"",)


def test_device_array_checks():
    """The Python handle refuses tensors the library would run past (host tensors, wrong dtype,
    too few bytes) before any launch; raw int addresses pass through."""
    import pytest
    import torch

    from salmon2_amd.hip import _need

    _need(None, 8, "x")
    _need(12345, 8, "x")
    with pytest.raises(ValueError, match="GPU tensor"):
        _need(torch.zeros(4, dtype=torch.float64), 8, "x")
    with pytest.raises(ValueError, match="GPU tensor"):
        _need(torch.zeros((4, 4), dtype=torch.float64).t(), 8, "x")
