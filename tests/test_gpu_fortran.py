"""The Fortran drop-in (iso_c_binding wrappers with SALMON's signatures) on the GPU, through
its test driver tests/fortran/test_dropin.f90:

* check: update_kvector_nonlocalpt_hip, hpsi_hip, taylor_hip, calc_density_hip and
  calc_current_hip on SALMON-type host arrays against the oracle (1e-12 / 1e-9), and the same
  Taylor step with the orbitals resident (salmon_hip_enter_data; no host copy at the call)
  bitwise equal to the host-array one;
* rt: the fixed-potential RT loop driven from Fortran with spsi_in, spsi_out and tpsi resident
  -- J(t) against the oracle's run (1e-9) and bitwise equal to the Python-hosted run of the
  same library;
* scf: hartree_hip, exchange_correlation_hip (PZ, nspin = 1 and 2) and update_vlocal_hip on
  SALMON-type host arrays, then taylor_hip with the new V_local, against the oracle (1e-12).
"""
import dataclasses
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


@pytest.mark.parametrize("name", ["c1", "spin2"])
def test_fortran_potential_wrappers(oracle_lib, tmp_path, name):
    s = CHECK_CASES[name]()
    A = np.array([0.006, 0.013, -0.009])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    io.write_input(fin, s, A)
    print(f"[gpu] fortran drop-in potential wrappers case={name}", flush=True)
    r = subprocess.run([_exe(), "scf", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = io.read_scf(fout, s)
    O = oracle_lib.Oracle(s)
    rho = O.density(s.psi)
    rhot = rho[0] + rho[1] if s.nspin == 2 else rho[0]
    vh = oracle_lib.hartree(rhot, s.al)
    if s.nspin == 2:
        vxc, exc = oracle_lib.xc_pz_sp(rho)
        trho = rho[0] + rho[1] + 1e-20
    else:
        v1, exc = oracle_lib.xc_pz(rho[0])
        vxc = v1[None]
        trho = rho[0]
    e_xc = float((exc * trho).sum() * s.hvol)
    vloc = s.vlocal[0][None] + vh[None] + vxc
    O2 = oracle_lib.Oracle(dataclasses.replace(s, vlocal=np.ascontiguousarray(vloc)))
    O2.update_kvector(A)
    tout, _, _ = O2.taylor(A, s.psi)
    ny = s.ny
    errs = dict(vh=rel(d["vh"], vh), vxc=rel(d["vxc"], vxc), e_xc=abs(d["exc_total"] - e_xc) / abs(e_xc),
                vlocal=rel(d["vlocal"], vloc), taylor=rel(d["tout"][..., :ny, :], tout[..., :ny, :]))
    print(f"[parity] fortran potential {name}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= 1e-12, k
