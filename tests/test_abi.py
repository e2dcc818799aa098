"""The C-ABI library loads without a GPU and exports every function include/salmon_hip.h declares."""
import ctypes
import os
import re

from salmon2_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "salmon_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(salmon_hip_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_binding_list():
    assert set(declared()) == set(hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hip.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially():
    assert hip.load().salmon_hip_last_error() in (b"",)
