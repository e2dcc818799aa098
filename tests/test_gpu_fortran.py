"""The Fortran drop-in (iso_c_binding wrappers with SALMON's signatures) on the GPU, through
its test driver tests/fortran/test_dropin.f90:

* check: update_kvector_nonlocalpt_hip, hpsi_hip, taylor_hip, calc_density_hip and
  calc_current_hip on SALMON-type host arrays against the oracle (1e-12 / 1e-9), and the same
  Taylor step with the orbitals resident (salmon_hip_enter_data; no host copy at the call)
  bitwise equal to the host-array one;
* rt: the fixed-potential RT loop driven from Fortran with spsi_in, spsi_out and tpsi resident
  -- J(t) against the oracle's run (1e-9) and bitwise equal to the Python-hosted run of the
  same library.
"""
import os
import subprocess

import numpy as np
import pytest

from salmon2_amd import build, system
from tests.fortran import dropin_io as io

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _exe():
    assert os.path.exists(build.DRIVER_EXE), "tests/fortran/_build/test_dropin not built"
    return build.DRIVER_EXE


CHECK_CASES = {
    "tiny": lambda: system.config("tiny"),
    "c1": lambda: system.config("c1"),
    # spin-polarised (V_local(ispin) per channel) and s+p+d projectors through the same wrappers
    "spin2": lambda: system.make_si_system(n=(16, 14, 12), no=3, nkgrid=(1, 1, 2), seed=47, nspin=2),
    "spd": lambda: system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), seed=53, lmax=2),
}


@pytest.mark.parametrize("name", list(CHECK_CASES))
def test_fortran_dropin_matches_oracle(oracle_lib, tmp_path, name):
    s = CHECK_CASES[name]()
    A = np.array([0.011, -0.007, 0.004])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    io.write_input(fin, s, A)
    print(f"[gpu] fortran drop-in check case={name}", flush=True)
    r = subprocess.run([_exe(), "check", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = io.read_check(fout, s)
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ny = s.ny
    e_h = rel(d["hpsi"][..., :ny, :], O.hpsi(A, s.psi)[..., :ny, :])
    ref_out, _, _ = O.taylor(A, s.psi)
    e_t = rel(d["tout"][..., :ny, :], ref_out[..., :ny, :])
    e_r = rel(d["rho"], O.density(ref_out))
    e_c = rel(d["curr"], O.current(A, ref_out))
    print(f"[parity] fortran {name}: hpsi {e_h:.2e} taylor {e_t:.2e} rho {e_r:.2e} curr {e_c:.2e}", flush=True)
    assert e_h <= 1e-12 and e_t <= 1e-12 and e_r <= 1e-12 and e_c <= 1e-9
    assert np.array_equal(d["tout_resident"], d["tout"])


def test_fortran_resident_rt_loop(oracle_lib, tmp_path):
    import torch

    from salmon2_amd.rt import RTPropagator

    nt = 20
    s = system.config("c1", nt=nt)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    io.write_input(fin, s, np.zeros(3), nt=nt, nwarm=2)
    print(f"[gpu] fortran drop-in resident RT loop, c1, {nt} steps", flush=True)
    r = subprocess.run([_exe(), "rt", str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = io.read_rt(fout, s, nt)
    _, cur, _, _ = oracle_lib.Oracle(s).rt_run(nt=nt)
    e = rel(d["curr"], cur)
    p = RTPropagator(s)
    py = p.run(nt)
    p.close()
    torch.cuda.empty_cache()
    print(f"[parity] fortran resident RT loop J(t) vs oracle {e:.2e}; vs python-hosted max diff "
          f"{np.abs(d['curr'] - py['curr']).max():.2e}", flush=True)
    assert e <= 1e-9
    assert np.array_equal(d["curr"], py["curr"])
