"""The Fortran drop-in (iso_c_binding wrappers with SALMON's signatures) on the GPU.

Runs salmon2_amd/fortran/test_dropin, which calls update_kvector_nonlocalpt_hip,
hpsi_hip, taylor_hip, calc_density_hip and calc_current_hip on SALMON-type
arguments, and checks every output against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from salmon2_amd import build, system

pytestmark = pytest.mark.gpu


def write_input(path, s, vec_ac):
    zc = s.zc()
    with open(path, "wb") as f:
        np.array([s.nx, s.ny, s.nz, s.ly, s.nspin, s.no, s.nk, s.natom, s.nps, s.nlma, len(zc)],
                 dtype=np.int32).tofile(f)
        np.array(list(s.hgs) + [s.hvol, s.coef_lap0], dtype=np.float64).tofile(f)
        np.ascontiguousarray(s.coef_lap, np.float64).tofile(f)
        np.ascontiguousarray(s.coef_nab, np.float64).tofile(f)
        np.asarray(vec_ac, np.float64).tofile(f)
        for a in (s.vec_k, s.wtk, s.rocc, s.vlocal):
            np.ascontiguousarray(a, np.float64).tofile(f)
        for a in (s.mps, s.jxyz, s.ia_tbl):
            np.ascontiguousarray(a, np.int32).tofile(f)
        for a in (s.rxyz, s.rinv_uvu, s.uv):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(zc, np.complex128).tofile(f)
        np.ascontiguousarray(s.psi, np.complex128).tofile(f)


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_fortran_dropin_matches_oracle(oracle_lib, tmp_path, name):
    exe = build.FORTRAN_EXE
    assert os.path.exists(exe), "salmon2_amd/fortran/test_dropin not built"
    s = system.config(name)
    A = np.array([0.011, -0.007, 0.004])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, s, A)
    print(f"[gpu] fortran drop-in case={name}", flush=True)
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    shp = s.psi.shape
    n = s.psi.size
    g = s.nx * s.ny * s.nz
    raw = np.fromfile(fout, dtype=np.uint8)
    off = 0

    def take(count, dt):
        nonlocal off
        nb = count * np.dtype(dt).itemsize
        out = raw[off:off + nb].view(dt)
        off += nb
        return out

    hpsi = take(n, np.complex128).reshape(shp)
    tout = take(n, np.complex128).reshape(shp)
    rho = take(g * s.nspin, np.float64).reshape(s.nspin, s.nz, s.ny, s.nx)
    curr = take(3 * s.nspin, np.float64).reshape(s.nspin, 3)
    assert off == raw.size
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ny = s.ny
    e_h = rel(hpsi[..., :ny, :], O.hpsi(A, s.psi)[..., :ny, :])
    ref_out, _, _ = O.taylor(A, s.psi)
    e_t = rel(tout[..., :ny, :], ref_out[..., :ny, :])
    e_r = rel(rho, O.density(ref_out))
    e_c = rel(curr, O.current(A, ref_out))
    print(f"[parity] fortran {name}: hpsi {e_h:.2e} taylor {e_t:.2e} rho {e_r:.2e} curr {e_c:.2e}", flush=True)
    assert e_h <= 1e-12 and e_t <= 1e-12 and e_r <= 1e-12 and e_c <= 1e-9
