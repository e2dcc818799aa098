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
