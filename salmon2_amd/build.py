"""Build the gfx950 HIP library in-tree (salmon2_amd/libsalmon_hip.so)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libsalmon_hip.so")

# All HIP code is built with packed-FP32 instruction generation disabled.
HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build_hip(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, "salmon_hip.hip"), os.path.join(CSRC, "salmon_unet.hip"),
            os.path.join(CSRC, "salmon_pp.cpp")]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, LIB)
    return LIB


FDIR = os.path.join(PKG, "fortran")
FBUILD = os.path.join(FDIR, "_build")
FORTRAN_SRCS = ["salmon_hip_mod.f90", "salmon_hip_structures_subset.f90", "salmon_hip_dropin.f90"]
# the drop-in's test driver (test infrastructure, built beside the product objects)
DRIVER_DIR = os.path.join(ROOT, "tests", "fortran")
DRIVER_SRC = os.path.join(DRIVER_DIR, "test_dropin.f90")
DRIVER_EXE = os.path.join(DRIVER_DIR, "_build", "test_dropin")


def flang() -> str:
    for cand in (os.environ.get("FC_FLANG"), "/opt/rocm/lib/llvm/bin/flang", "flang"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "flang"


def _flang(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build_fortran(verbose: bool = False) -> str:
    """iso_c_binding layer + SALMON-signature wrappers (objects and modules in fortran/_build),
    and the wrappers' test driver linked to libsalmon_hip.so."""
    os.makedirs(FBUILD, exist_ok=True)
    objs = []
    for src in FORTRAN_SRCS:
        obj = os.path.join(FBUILD, src.replace(".f90", ".o"))
        _flang([flang(), "-O2", "-fPIC", "-module-dir", FBUILD, "-c", os.path.join(FDIR, src), "-o", obj])
        objs.append(obj)
    if not os.path.exists(DRIVER_SRC):
        return FBUILD
    dbuild = os.path.dirname(DRIVER_EXE)
    os.makedirs(dbuild, exist_ok=True)
    dobj = os.path.join(dbuild, "test_dropin.o")
    _flang([flang(), "-O2", "-fPIC", "-module-dir", dbuild, "-I", FBUILD, "-c", DRIVER_SRC, "-o", dobj])
    rpath = os.path.relpath(PKG, dbuild)
    _flang([flang(), "-o", DRIVER_EXE, *objs, dobj, "-L", PKG, "-lsalmon_hip", f"-Wl,-rpath,$ORIGIN/{rpath}"])
    return DRIVER_EXE


if __name__ == "__main__":
    print(build_hip(verbose=True))
    print(build_fortran(verbose=True))
