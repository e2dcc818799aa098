"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (liboracle.so).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
only.  It is the checker and the reported CPU baseline, never the thing
measured or shipped.  See salmon_oracle.c for the reference loops each
routine restates and what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libzstencil_ref.so")
REF_XC_LIB = os.path.join(HERE, "_ref", "libxc_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSystem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ly", C.c_int),
        ("nspin", C.c_int), ("no", C.c_int), ("nk", C.c_int),
        ("coef_lap", C.c_double * 12), ("coef_nab", C.c_double * 12),
        ("coef_lap0", C.c_double), ("hgs", C.c_double * 3), ("hvol", C.c_double),
        ("stencil_mode", C.c_int),
        ("vec_k", _dp), ("rocc", _dp), ("wtk", _dp), ("vlocal", _dp),
        ("natom", C.c_int), ("nps", C.c_int), ("nlma", C.c_int),
        ("mps", _ip), ("jxyz", _ip), ("rxyz", _dp), ("ia_tbl", _ip),
        ("rinv_uvu", _dp), ("uv", _dp), ("zekr", _dp),
    ]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and the reference stencil when its sources exist)."""
    kw = dict(cwd=HERE, check=True)
    if quiet:
        kw.update(stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    subprocess.run(["make", "-s", "liboracle.so"], **kw)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "ref"], **kw)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_zstencil_seq.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_zstencil_omp.argtypes = L.orc_zstencil_seq.argtypes
        L.orc_stencil_coefs.argtypes = [_dp, _dp, _dp, _dp]
        S = C.POINTER(OrcSystem)
        L.orc_update_kvector.argtypes = [S, _dp]
        L.orc_zpseudo.argtypes = [S, _dp, _dp]
        L.orc_hpsi.argtypes = [S, _dp, _dp, _dp]
        L.orc_taylor.argtypes = [S, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_taylor_coefs.argtypes = [C.c_double, C.c_int, C.c_int, _dp]
        L.orc_density.argtypes = [S, _dp, _dp]
        L.orc_current.argtypes = [S, _dp, _dp, _dp]
        L.orc_dipole.argtypes = [S, _dp, _dp]
        L.orc_ac_asin2cos.argtypes = [C.c_double] * 4 + [_dp, C.c_double, C.c_int, _dp]
        L.orc_rt_run.argtypes = [S, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_rt_run.restype = C.c_int
        L.orc_hartree.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_xc_pz.argtypes = [C.c_long, _dp, _dp, _dp]
        L.orc_xc_pz_sp.argtypes = [C.c_long, _dp, _dp, _dp]
        L.orc_rt_run_scf.argtypes = [S, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                     _dp, _dp]
        L.orc_rt_run_scf.restype = C.c_int
        L.orc_rt_run_full.argtypes = [S, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                      _dp, _dp, _dp, _dp, _dp]
        L.orc_rt_run_full.restype = C.c_int
        L.orc_rt_run_prop.argtypes = (L.orc_rt_run_full.argtypes[:4] + [C.c_int]
                                      + L.orc_rt_run_full.argtypes[4:])
        L.orc_rt_run_prop.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """Holds an orc_system view of a salmon2_amd.system.System (arrays kept alive here)."""

    def __init__(self, sysm, stencil_mode: int = -1):
        self.s = sysm
        self._keep = []

        def keep(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a

        o = OrcSystem()
        o.nx, o.ny, o.nz, o.ly = sysm.nx, sysm.ny, sysm.nz, sysm.ly
        o.nspin, o.no, o.nk = sysm.nspin, sysm.no, sysm.nk
        for i, v in enumerate(np.asarray(sysm.coef_lap).ravel()):
            o.coef_lap[i] = v
        for i, v in enumerate(np.asarray(sysm.coef_nab).ravel()):
            o.coef_nab[i] = v
        o.coef_lap0 = sysm.coef_lap0
        for d in range(3):
            o.hgs[d] = sysm.hgs[d]
        o.hvol = sysm.hvol
        o.stencil_mode = stencil_mode
        o.vec_k = _p(keep(sysm.vec_k, np.float64))
        o.rocc = _p(keep(sysm.rocc, np.float64))
        o.wtk = _p(keep(sysm.wtk, np.float64))
        o.vlocal = _p(keep(sysm.vlocal, np.float64))
        o.natom, o.nps, o.nlma = sysm.natom, sysm.nps, sysm.nlma
        o.mps = _p(keep(sysm.mps, np.int32), _ip)
        o.jxyz = _p(keep(sysm.jxyz, np.int32), _ip)
        o.rxyz = _p(keep(sysm.rxyz, np.float64))
        o.ia_tbl = _p(keep(sysm.ia_tbl, np.int32), _ip)
        o.rinv_uvu = _p(keep(sysm.rinv_uvu, np.float64))
        o.uv = _p(keep(sysm.uv, np.float64))
        self.zekr = np.zeros((sysm.nk, max(sysm.nlma, 1), max(sysm.nps, 1)), dtype=np.complex128)
        o.zekr = _p(self.zekr.view(np.float64))
        self.o = o

    @property
    def ref(self):
        return C.byref(self.o)

    def update_kvector(self, vec_ac):
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        if self.s.natom:
            lib().orc_update_kvector(self.ref, _p(a))

    def hpsi(self, vec_ac, tpsi):
        """H tpsi; the phase table must be current (update_kvector)."""
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        t = np.ascontiguousarray(tpsi, dtype=np.complex128)
        h = np.zeros_like(t)
        lib().orc_hpsi(self.ref, _p(a), _p(t.view(np.float64)), _p(h.view(np.float64)))
        return h

    def zpseudo(self, tpsi, htpsi):
        t = np.ascontiguousarray(tpsi, dtype=np.complex128)
        h = np.array(htpsi, dtype=np.complex128, copy=True)
        lib().orc_zpseudo(self.ref, _p(t.view(np.float64)), _p(h.view(np.float64)))
        return h

    def taylor(self, vec_ac, psi_in, zc=None):
        """Returns (psi_out, clobbered tspsi_in, sshtpsi) exactly as taylor.f90 leaves them."""
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        zc = self.s.zc() if zc is None else np.asarray(zc, dtype=np.complex128)
        zc = np.ascontiguousarray(zc)
        tin = np.array(psi_in, dtype=np.complex128, copy=True)
        out = np.zeros_like(tin)
        w = np.zeros_like(tin)
        lib().orc_taylor(self.ref, _p(a), _p(zc.view(np.float64)), len(zc), _p(tin.view(np.float64)),
                         _p(out.view(np.float64)), _p(w.view(np.float64)))
        return out, tin, w

    def density(self, psi):
        p = np.ascontiguousarray(psi, dtype=np.complex128)
        rho = np.zeros((self.s.nspin, self.s.nz, self.s.ny, self.s.nx))
        lib().orc_density(self.ref, _p(p.view(np.float64)), _p(rho))
        return rho

    def current(self, vec_ac, psi):
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        p = np.ascontiguousarray(psi, dtype=np.complex128)
        c = np.zeros((self.s.nspin, 3))
        lib().orc_current(self.ref, _p(a), _p(p.view(np.float64)), _p(c))
        return c

    def dipole(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        out = np.zeros(4)
        lib().orc_dipole(self.ref, _p(r), _p(out))
        return out

    def _ac(self, nt, ac):
        """The field table rows 0 .. nt+1: the given one (e.g. the rows of a continued run), else
        the system's pulse from t = 0."""
        s = self.s
        if ac is not None:
            ac = np.ascontiguousarray(ac, dtype=np.float64)
            if ac.shape != (nt + 2, 3):
                raise ValueError(f"field table of shape {ac.shape}, expected {(nt + 2, 3)}")
            return ac
        return np.ascontiguousarray(s.pulse.table(s.dt, nt) if s.pulse else np.zeros((nt + 2, 3)))

    def rt_run(self, nt=None, psi=None, ac=None):
        """Plain RT loop (fixed V_local, middlepoint, 'tr'); returns final psi, J(t), dipole(t), rNe(t)."""
        s = self.s
        nt = s.nt if nt is None else nt
        ac = self._ac(nt, ac)
        p = np.array(s.psi if psi is None else psi, dtype=np.complex128, copy=True)
        cur = np.zeros((nt, 3))
        dip = np.zeros((nt, 3))
        rne = np.zeros(nt)
        rc = lib().orc_rt_run(self.ref, s.dt, s.n_hamil, nt, _p(ac), _p(p.view(np.float64)), _p(cur),
                              _p(dip), _p(rne))
        if rc != 0:
            raise MemoryError("oracle rt_run allocation failed")
        return p, cur, dip, rne


    def rt_run_scf(self, pc: bool, nt=None, psi=None, ac=None):
        """RT loop with V_local = Vpsl + Vh[rho] + Vxc[rho] rebuilt every step (system.vlocal is Vpsl),
        optionally with the predictor-corrector; returns final psi, J(t), dipole(t), rNe(t), final V_local."""
        s = self.s
        nt = s.nt if nt is None else nt
        ac = self._ac(nt, ac)
        p = np.array(s.psi if psi is None else psi, dtype=np.complex128, copy=True)
        vpsl = np.ascontiguousarray(s.vlocal[0], dtype=np.float64)
        vloc = np.zeros((s.nspin,) + vpsl.shape) if s.nspin == 2 else np.zeros_like(vpsl)
        al = np.ascontiguousarray(s.al, dtype=np.float64)
        cur = np.zeros((nt, 3))
        dip = np.zeros((nt, 3))
        rne = np.zeros(nt)
        saved = self.o.vlocal
        try:
            rc = lib().orc_rt_run_scf(self.ref, _p(vpsl), _p(vloc), _p(al), int(bool(pc)), s.dt, s.n_hamil, nt,
                                      _p(ac), _p(p.view(np.float64)), _p(cur), _p(dip), _p(rne))
        finally:
            self.o.vlocal = saved
        if rc != 0:
            raise RuntimeError(f"oracle rt_run_scf failed ({rc})")
        return p, cur, dip, rne, vloc


    def rt_run_full(self, scf=False, pc=False, trans_longi="tr", nt=None, psi=None, propagator="middlepoint"):
        """General loop ('middlepoint' or 'aetrs'); returns final psi, J(t), dipole(t), rNe(t),
        Ac_tot (nt+2, 3), final V_local."""
        s = self.s
        nt = s.nt if nt is None else nt
        ac = np.ascontiguousarray(s.pulse.table(s.dt, nt) if s.pulse else np.zeros((nt + 2, 3)))
        p = np.array(s.psi if psi is None else psi, dtype=np.complex128, copy=True)
        vpsl = np.ascontiguousarray(s.vlocal[0], dtype=np.float64)
        vloc = np.zeros((s.nspin,) + vpsl.shape) if s.nspin == 2 else np.zeros_like(vpsl)
        al = np.ascontiguousarray(s.al, dtype=np.float64)
        cur = np.zeros((nt, 3))
        dip = np.zeros((nt, 3))
        rne = np.zeros(nt)
        if trans_longi not in ("tr", "lo"):
            raise ValueError(trans_longi)
        if propagator not in ("middlepoint", "aetrs"):
            raise ValueError(propagator)
        saved = self.o.vlocal
        try:
            rc = lib().orc_rt_run_prop(self.ref, int(bool(scf)), int(bool(pc)), int(trans_longi == "lo"),
                                       int(propagator == "aetrs"), _p(vpsl), _p(vloc), _p(al), s.dt, s.n_hamil, nt,
                                       _p(ac), _p(p.view(np.float64)), _p(cur), _p(dip), _p(rne))
        finally:
            self.o.vlocal = saved
        if rc != 0:
            raise RuntimeError(f"oracle rt_run_full failed ({rc})")
        return p, cur, dip, rne, ac, (vloc if scf else None)


def hartree(rho, al):
    """Hartree potential of a periodic orthogonal cell; rho is (nz, ny, nx)."""
    r = np.ascontiguousarray(rho, dtype=np.float64)
    nz, ny, nx = r.shape
    v = np.zeros_like(r)
    lib().orc_hartree(nx, ny, nz, _p(np.ascontiguousarray(al, dtype=np.float64)), _p(r), _p(v))
    return v


def xc_pz(rho):
    """PZ81 LDA (unpolarised): returns (vxc, exc) per point."""
    r = np.ascontiguousarray(rho, dtype=np.float64)
    v = np.zeros_like(r)
    e = np.zeros_like(r)
    lib().orc_xc_pz(r.size, _p(r), _p(v), _p(e))
    return v, e


def xc_pz_sp(rho_s):
    """PZ81 LDA, spin-polarised: rho_s is (2, ...) (the two channel densities); returns
    (vxc_s (2, ...), exc (...))."""
    r = np.ascontiguousarray(rho_s, dtype=np.float64)
    if r.shape[0] != 2:
        raise ValueError("rho_s: two spin channels along the first axis")
    v = np.zeros_like(r)
    e = np.zeros(r.shape[1:])
    lib().orc_xc_pz_sp(r[0].size, _p(r), _p(v), _p(e))
    return v, e


def zstencil(nx, ny, nz, ly, tpsi, V, lap0, lapt, nabt, mode="seq"):
    """Oracle stencil on one orbital block (nz, ly, nx)."""
    t = np.ascontiguousarray(tpsi, dtype=np.complex128)
    h = np.zeros_like(t)
    Vv = np.ascontiguousarray(V, dtype=np.float64)
    lt = np.ascontiguousarray(lapt, dtype=np.float64).ravel()
    nt_ = np.ascontiguousarray(nabt, dtype=np.float64).ravel()
    f = lib().orc_zstencil_seq if mode == "seq" else lib().orc_zstencil_omp
    f(nx, ny, nz, ly, _p(t.view(np.float64)), _p(h.view(np.float64)), _p(Vv), lap0, _p(lt), _p(nt_))
    return h


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_zstencil(nx, ny, nz, ly, tpsi, V, lap0, lapt, nabt, mode="seq"):
    """SALMON's own zstencil_typical_{seq,omp}, compiled from the reference sources.

    Called with SALMON's periodic single-process conventions: is = 1, ie = n,
    is_array = is, ie_array(2) = ny + 1 (padding), idx(j) = mod(j+n-1, n)+1
    over is-4 .. ie+4 (initialization.f90:574-609)."""
    L = C.CDLL(REF_LIB)
    t = np.ascontiguousarray(tpsi, dtype=np.complex128)
    h = np.zeros_like(t)
    Vv = np.ascontiguousarray(V, dtype=np.float64)
    i4 = lambda *v: (C.c_int * len(v))(*v)
    is_ = i4(1, 1, 1)
    ie = i4(nx, ny, nz)
    is_arr = i4(1, 1, 1)
    ie_arr = i4(nx, ly, nz)
    tabs = [i4(*[((j + n - 1) % n) + 1 for j in range(1 - 4, n + 4 + 1)]) for n in (nx, ny, nz)]
    lap0_ = C.c_double(lap0)
    lt = np.ascontiguousarray(lapt, dtype=np.float64).ravel()
    nb = np.ascontiguousarray(nabt, dtype=np.float64).ravel()
    if mode == "seq":
        L.zstencil_typical_seq_(is_arr, ie_arr, is_, ie, tabs[0], tabs[1], tabs[2], is_, ie,
                                _p(t.view(np.float64)), _p(h.view(np.float64)), _p(Vv), C.byref(lap0_),
                                _p(lt), _p(nb))
    else:
        L.zstencil_typical_omp_(is_arr, ie_arr, is_, ie, tabs[0], tabs[1], tabs[2],
                                _p(t.view(np.float64)), _p(h.view(np.float64)), _p(Vv), C.byref(lap0_),
                                _p(lt), _p(nb))
    return h


def ref_xc_available() -> bool:
    return os.path.exists(REF_XC_LIB)


def ref_xc_pz(rho):
    """SALMON's exc_cor_pz (builtin_pz.f90:23-51) compiled from the reference source, called as
    exec_builtin_pz does for nspin = 1 (rho_s = rho * 0.5, salmon_xc.f90:778, 795); returns
    (vxc, exc)."""
    L = C.CDLL(REF_XC_LIB)
    r = np.ascontiguousarray(rho, dtype=np.float64)
    rs = np.ascontiguousarray(r * 0.5)
    v, e, ee = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
    L.ref_exc_cor_pz(C.c_int(r.size), _p(rs), _p(e), _p(ee), _p(v))
    return v, e


def ref_xc_pz_sp(rho_s):
    """SALMON's exc_cor_pz_sp (builtin_pz_sp.f90:23-40) compiled from the reference source;
    rho_s is (2, ...); returns (vxc_s (2, ...), exc)."""
    L = C.CDLL(REF_XC_LIB)
    r = np.ascontiguousarray(rho_s, dtype=np.float64)
    v = np.zeros_like(r)
    e, ee = np.zeros(r.shape[1:]), np.zeros(r.shape[1:])
    L.ref_exc_cor_pz_sp(C.c_int(r[0].size), _p(r), _p(e), _p(ee), _p(v))
    return v, e
