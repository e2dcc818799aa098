"""Input and output files of the Fortran drop-in driver tests/fortran/test_dropin.f90
(test infrastructure; stream files in the driver's read order)."""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
EXE = os.path.join(BUILD, "test_dropin")
SRC = os.path.join(HERE, "test_dropin.f90")


def write_input(path, s, vec_ac, psi=None, nt=0, nwarm=0):
    """System s (salmon2_amd.system.System; orbitals `psi` or s.psi), the field vec_ac of the
    single-call checks, for the RT loop nt steps (nwarm untimed) with s's pulse table, and the
    cell lengths (system%primitive_a of an orthogonal cell)."""
    zc = s.zc()
    psi = s.psi if psi is None else psi
    ac = s.pulse.table(s.dt, nt) if (s.pulse is not None and nt > 0) else np.zeros((nt + 2, 3))
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
        np.ascontiguousarray(psi, np.complex128).tofile(f)
        np.array([nt, nwarm], dtype=np.int32).tofile(f)
        np.ascontiguousarray(ac, np.float64).tofile(f)
        np.ascontiguousarray(s.al, np.float64).tofile(f)


class _Reader:
    def __init__(self, path):
        self.raw = np.fromfile(path, dtype=np.uint8)
        self.off = 0

    def take(self, count, dt):
        nb = count * np.dtype(dt).itemsize
        out = self.raw[self.off:self.off + nb].view(dt)
        self.off += nb
        return out

    def done(self):
        return self.off == self.raw.size


def read_check(path, s):
    r = _Reader(path)
    shp, n, g = s.psi.shape, s.psi.size, s.nx * s.ny * s.nz
    out = dict(hpsi=r.take(n, np.complex128).reshape(shp), tout=r.take(n, np.complex128).reshape(shp),
               rho=r.take(g * s.nspin, np.float64).reshape(s.nspin, s.nz, s.ny, s.nx),
               curr=r.take(3 * s.nspin, np.float64).reshape(s.nspin, 3),
               tout_resident=r.take(n, np.complex128).reshape(shp))
    assert r.done()
    return out


def read_rt(path, s, nt):
    r = _Reader(path)
    el = float(r.take(1, np.float64)[0])
    jt = r.take(3 * nt, np.float64).reshape(nt, 3)
    n = s.nk * s.no * s.nspin * s.nz * s.ly * s.nx
    psi = r.take(n, np.complex128).reshape(s.nk, s.no, s.nspin, s.nz, s.ly, s.nx)
    assert r.done()
    return dict(elapsed=el, curr=jt, psi=psi)


def read_scf(path, s):
    r = _Reader(path)
    g = s.nx * s.ny * s.nz
    shp3 = (s.nz, s.ny, s.nx)
    out = dict(vh=r.take(g, np.float64).reshape(shp3),
               vxc=r.take(g * s.nspin, np.float64).reshape((s.nspin,) + shp3),
               exc_total=float(r.take(1, np.float64)[0]),
               vlocal=r.take(g * s.nspin, np.float64).reshape((s.nspin,) + shp3),
               tout=r.take(s.psi.size, np.complex128).reshape(s.psi.shape))
    assert r.done()
    return out
