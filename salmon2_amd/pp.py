"""ctypes host interface of the projector-table builder (include/salmon_pp.h).

Reads KY-format pseudopotentials ("*_rps.dat") and builds, for an orthogonal periodic
cell, the tables the device path takes: sphere points, uv and rinv_uvu for
salmon_hip_set_nonlocal, and the local pseudopotential Vpsl (SALMON's init_ps,
src/atom/pp/prep_pp.f90:22-527, and input_pp, src/atom/pp/input_pp.f90:28-341).
Arrays come back in the axis-reversed C order used across this package
(``uv[lma, j]`` is SALMON's ``uv(j, lma)``, ``vpsl[iz, iy, ix]``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import hip

EXPORTS = ("salmon_pp_create", "salmon_pp_destroy", "salmon_pp_element", "salmon_pp_sizes", "salmon_pp_nonlocal",
           "salmon_pp_vpsl")

_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_bound = False


def _lib():
    global _bound
    L = hip.load()
    if not _bound:
        L.salmon_pp_create.argtypes = [C.c_int, C.POINTER(C.c_char_p), _ip, _ip, C.POINTER(C.c_void_p)]
        L.salmon_pp_destroy.argtypes = [C.c_void_p]
        L.salmon_pp_element.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]
        cell = [C.c_void_p, C.c_int, _ip, _dp, _dp, _ip]
        L.salmon_pp_sizes.argtypes = cell + [C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.salmon_pp_nonlocal.argtypes = cell + [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _dp, _dp]
        L.salmon_pp_vpsl.argtypes = cell + [_dp]
        _bound = True
    return L


def _chk(rc, what):
    if rc != 0:
        raise hip.SalmonHipError(f"{what}: {hip.load().salmon_hip_last_error().decode()}")


class PseudoPotential:
    """Element tables of KY files (one per element) with SALMON's Lmax_ps / Lloc_ps."""

    def __init__(self, files, lloc_ps, lmax_ps=None):
        L = _lib()
        n = len(files)
        arr = (C.c_char_p * n)(*[os.fsencode(f) for f in files])
        lmax = np.array(lmax_ps if lmax_ps is not None else [-1] * n, dtype=np.int32)
        lloc = np.array(lloc_ps, dtype=np.int32)
        h = C.c_void_p()
        _chk(L.salmon_pp_create(n, arr, lmax, lloc, C.byref(h)), "salmon_pp_create")
        self.h = h
        self.nelem = n

    def close(self):
        if self.h:
            _lib().salmon_pp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def element(self, ielem=1):
        zps, mlps, lref, nrps, nlma = (C.c_int() for _ in range(5))
        rps = C.c_double()
        _chk(_lib().salmon_pp_element(self.h, ielem, C.byref(zps), C.byref(mlps), C.byref(lref), C.byref(rps),
                                      C.byref(nrps), C.byref(nlma)), "salmon_pp_element")
        return dict(zps=zps.value, mlps=mlps.value, lref=lref.value, rps=rps.value, nrps=nrps.value,
                    nlma_atom=nlma.value)

    @staticmethod
    def _cell(kion, rion, al, n):
        return (np.ascontiguousarray(kion, dtype=np.int32), np.ascontiguousarray(rion, dtype=np.float64),
                np.ascontiguousarray(al, dtype=np.float64), np.ascontiguousarray(n, dtype=np.int32))

    def nonlocal_tables(self, kion, rion, al, n):
        """kion (1-based element per atom), rion[natom, 3], al[3], n[3] ->
        (mps, nps, jxyz[natom, nps, 3] 1-based, rxyz[natom, nps, 3], nlma, ia_tbl, rinv_uvu, uv[nlma, nps])."""
        L = _lib()
        k, r, a, nn = self._cell(kion, rion, al, n)
        natom = len(k)
        nps, nlma = C.c_int(), C.c_int()
        _chk(L.salmon_pp_sizes(self.h, natom, k, r, a, nn, C.byref(nps), C.byref(nlma)), "salmon_pp_sizes")
        nps, nlma = nps.value, nlma.value
        mps = np.zeros(natom, dtype=np.int32)
        jxyz = np.zeros((natom, nps, 3), dtype=np.int32)
        rxyz = np.zeros((natom, nps, 3))
        ia_tbl = np.zeros(nlma, dtype=np.int32)
        rinv = np.zeros(nlma)
        uv = np.zeros((nlma, nps))
        _chk(L.salmon_pp_nonlocal(self.h, natom, k, r, a, nn, nps, nlma, mps, jxyz, rxyz, ia_tbl, rinv, uv),
             "salmon_pp_nonlocal")
        return mps, nps, jxyz, rxyz, nlma, ia_tbl, rinv, uv

    def vpsl(self, kion, rion, al, n):
        """Local pseudopotential on the grid, vpsl[iz, iy, ix]."""
        k, r, a, nn = self._cell(kion, rion, al, n)
        out = np.zeros((int(nn[2]), int(nn[1]), int(nn[0])))
        _chk(_lib().salmon_pp_vpsl(self.h, len(k), k, r, a, nn, out), "salmon_pp_vpsl")
        return out
