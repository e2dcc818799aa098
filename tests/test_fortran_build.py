"""The Fortran iso_c_binding layer, the SALMON-signature wrappers and their test driver
compile and link (no GPU needed)."""
import os
import shutil

import pytest

from salmon2_amd import build


@pytest.mark.skipif(shutil.which(build.flang()) is None and not os.path.exists(build.flang()),
                    reason="no Fortran compiler")
def test_fortran_layer_builds_and_links():
    exe = build.build_fortran()
    assert exe == build.DRIVER_EXE
    assert os.path.exists(exe) and os.access(exe, os.X_OK)
    for mod in ("salmon_hip_mod.mod", "salmon_hip_dropin.mod", "structures.mod"):
        assert os.path.exists(os.path.join(build.FBUILD, mod))


def test_fortran_bindings_name_declared_entry_points():
    """Every bind(c) name of the Fortran interface module is a function declared in
    include/salmon_hip.h (or a C library routine the module uses), with as many arguments."""
    import re

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    mod = open(os.path.join(root, "salmon2_amd", "fortran", "salmon_hip_mod.f90")).read()
    hdr = open(os.path.join(root, "include", "salmon_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    decls = {m.group(1): m.group(2) for m in re.finditer(r"\b(salmon_hip_\w+)\s*\(([^;()]*)\)\s*;", hdr)}
    found = re.findall(r"function\s+(\w+)\s*\(([^)]*)\)\s*&?\s*bind\(c,\s*name='(\w+)'\)", mod)
    assert len(found) >= 20
    for fname, args, cname in found:
        if not cname.startswith("salmon_hip_"):
            continue  # strlen
        assert fname == cname
        assert cname in decls, cname
        n_f = len([a for a in args.replace("&", " ").split(",") if a.strip()])
        n_c = len([a for a in decls[cname].split(",") if a.strip() and a.strip() != "void"])
        assert n_f == n_c, (cname, n_f, n_c)
