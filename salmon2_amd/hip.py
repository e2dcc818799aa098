"""ctypes binding of the C ABI in include/salmon_hip.h.

The library is the in-tree ``salmon2_amd/libsalmon_hip.so`` built by
``salmon2_amd.build``.  There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

EXPORTS = (
    "salmon_hip_last_error",
    "salmon_hip_create",
    "salmon_hip_destroy",
    "salmon_hip_set_stream",
    "salmon_hip_get_stream",
    "salmon_hip_set_nonlocal",
    "salmon_hip_set_vlocal",
    "salmon_hip_set_vlocal_device",
    "salmon_hip_update_kvector",
    "salmon_hip_set_vec_ac",
    "salmon_hip_hpsi",
    "salmon_hip_taylor",
    "salmon_hip_taylor_reuse",
    "salmon_hip_set_skip",
    "salmon_hip_copy_if",
    "salmon_hip_stream_copy",
    "salmon_hip_set_precision",
    "salmon_hip_autotune",
    "salmon_hip_set_launch_shape",
    "salmon_hip_set_overlap",
    "salmon_hip_batch_create",
    "salmon_hip_batch_destroy",
    "salmon_hip_batch_set_system",
    "salmon_hip_batch_prepare",
    "salmon_hip_batch_autotune",
    "salmon_hip_batch_set_time",
    "salmon_hip_batch_step",
    "salmon_hip_batch_capture",
    "salmon_hip_batch_replay",
    "salmon_hip_density",
    "salmon_hip_density_noalloc",
    "salmon_hip_current",
    "salmon_hip_current_next",
    "salmon_hip_dipole",
    "salmon_hip_set_poisson",
    "salmon_hip_hartree",
    "salmon_hip_xc_pz",
    "salmon_hip_update_vlocal",
    "salmon_hip_update_potential",
    "salmon_hip_vlocal_extrapolate",
    "salmon_hip_synchronize",
    "salmon_hip_prof_start",
    "salmon_hip_prof_stop",
    "salmon_hip_prof_end",
    "salmon_hip_malloc",
    "salmon_hip_free",
    "salmon_hip_memcpy_htod",
    "salmon_hip_memcpy_dtoh",
)

_lib = None


class SalmonHipError(RuntimeError):
    pass


def lib_path() -> str:
    return _build.LIB


def load():
    """Load libsalmon_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise SalmonHipError(f"{path} is missing: build it with `python -m salmon2_amd.build`")
    L = C.CDLL(path)
    L.salmon_hip_last_error.restype = C.c_char_p
    L.salmon_hip_create.argtypes = [C.c_int] * 7 + [_dp, _dp, C.c_double, _dp, _dp, _dp, _dp, C.c_double,
                                                   C.POINTER(C.c_void_p)]
    L.salmon_hip_destroy.argtypes = [C.c_void_p]
    L.salmon_hip_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.salmon_hip_get_stream.argtypes = [C.c_void_p]
    L.salmon_hip_get_stream.restype = C.c_void_p
    L.salmon_hip_set_nonlocal.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, _ip, _dp, _dp]
    L.salmon_hip_set_vlocal.argtypes = [C.c_void_p, _dp]
    L.salmon_hip_set_vlocal_device.argtypes = [C.c_void_p, C.c_void_p]
    L.salmon_hip_update_kvector.argtypes = [C.c_void_p, _dp]
    L.salmon_hip_set_vec_ac.argtypes = [C.c_void_p, _dp]
    L.salmon_hip_hpsi.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_taylor.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_taylor_reuse.argtypes = [C.c_void_p, _dp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_set_skip.argtypes = [C.c_void_p, C.c_void_p]
    L.salmon_hip_copy_if.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.salmon_hip_stream_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, _dp]
    L.salmon_hip_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.salmon_hip_autotune.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
    L.salmon_hip_set_launch_shape.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.salmon_hip_set_overlap.argtypes = [C.c_void_p, C.c_int]
    vp = C.c_void_p
    L.salmon_hip_batch_create.argtypes = [C.POINTER(vp), C.c_int, vp, C.POINTER(vp)]
    L.salmon_hip_batch_destroy.argtypes = [vp]
    L.salmon_hip_batch_set_system.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, vp]
    L.salmon_hip_batch_prepare.argtypes = [vp, _dp, C.c_int]
    L.salmon_hip_batch_autotune.argtypes = [vp, C.c_int, _ip, _dp]
    L.salmon_hip_batch_set_time.argtypes = [vp, C.c_int]
    L.salmon_hip_batch_step.argtypes = [vp]
    L.salmon_hip_batch_capture.argtypes = [vp]
    L.salmon_hip_batch_replay.argtypes = [vp, C.c_int]
    L.salmon_hip_density.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_density_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_current.argtypes = [C.c_void_p, _dp, C.c_void_p, C.c_void_p]
    L.salmon_hip_current_next.argtypes = [C.c_void_p, _dp, C.c_void_p, C.c_void_p, _dp]
    L.salmon_hip_dipole.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_set_poisson.argtypes = [C.c_void_p, _dp]
    L.salmon_hip_hartree.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_xc_pz.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.salmon_hip_update_vlocal.argtypes = [C.c_void_p] + [C.c_void_p] * 5
    L.salmon_hip_update_potential.argtypes = [C.c_void_p] + [C.c_void_p] * 4
    L.salmon_hip_vlocal_extrapolate.argtypes = [C.c_void_p] * 4
    L.salmon_hip_synchronize.argtypes = [C.c_void_p]
    L.salmon_hip_prof_start.argtypes = [C.c_void_p, C.c_int]
    L.salmon_hip_prof_stop.argtypes = [C.c_void_p, C.c_int, _dp, _ip]
    L.salmon_hip_prof_end.argtypes = [C.c_void_p]
    L.salmon_hip_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.salmon_hip_free.argtypes = [C.c_void_p]
    L.salmon_hip_memcpy_htod.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.salmon_hip_memcpy_dtoh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    _lib = L
    return L


def _chk(rc: int, what: str):
    if rc != 0:
        msg = load().salmon_hip_last_error().decode(errors="replace")
        raise SalmonHipError(f"{what} failed (code {rc}): {msg}")


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _ptr(t) -> int:
    """Device address of a torch tensor (or an int address)."""
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _need(t, nbytes, what, f64=True):
    """A torch tensor handed to the library as a device array must be a contiguous GPU tensor of
    at least nbytes (float64 where the library reads or writes doubles); raw int addresses are
    the caller's responsibility.  Raises ValueError instead of letting a kernel run past it."""
    if t is None or isinstance(t, int):
        return
    import torch

    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{what}: a contiguous GPU tensor is required")
    if f64 and t.dtype != torch.float64:
        raise ValueError(f"{what}: float64 expected, got {t.dtype}")
    if t.numel() * t.element_size() < nbytes:
        raise ValueError(f"{what}: {t.numel() * t.element_size()} bytes, at least {nbytes} needed")


class Handle:
    """One grid + orbital set on the GPU (wraps salmon_hip_create .. destroy)."""

    def __init__(self, sysm, stream=None):
        L = load()
        self.s = sysm
        h = C.c_void_p()
        lap = np.ascontiguousarray(sysm.coef_lap, dtype=np.float64).ravel()
        nab = np.ascontiguousarray(sysm.coef_nab, dtype=np.float64).ravel()
        veck = np.ascontiguousarray(sysm.vec_k, dtype=np.float64)
        rocc = np.ascontiguousarray(sysm.rocc, dtype=np.float64)
        wtk = np.ascontiguousarray(sysm.wtk, dtype=np.float64)
        hgs = np.ascontiguousarray(sysm.hgs, dtype=np.float64)
        _chk(L.salmon_hip_create(sysm.nx, sysm.ny, sysm.nz, sysm.ly, sysm.nspin, sysm.no, sysm.nk, _p(lap),
                                 _p(nab), float(sysm.coef_lap0), _p(veck), _p(rocc), _p(wtk), _p(hgs),
                                 float(sysm.hvol), C.byref(h)), "salmon_hip_create")
        self.h = h
        # bytes of one grid channel (double), of V_local / rho (nspin channels), of an orbital set
        self._g = 8 * sysm.nx * sysm.ny * sysm.nz
        self._gs = self._g * sysm.nspin
        self._orb = 16 * sysm.nk * sysm.no * sysm.nspin * sysm.nz * sysm.ly * sysm.nx
        if stream is not None:
            self.set_stream(stream)
        if sysm.natom > 0:
            self.set_nonlocal(sysm)
        self.set_vlocal(sysm.vlocal)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().salmon_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """stream: torch.cuda.Stream or raw hipStream_t int; torch's default stream (and 0 / None)
        is HIP's null stream, so library work stays ordered with torch work on it."""
        raw = getattr(stream, "cuda_stream", stream)
        _chk(load().salmon_hip_set_stream(self.h, C.c_void_p(raw or None)), "set_stream")

    def set_nonlocal(self, s):
        self._nl = [np.ascontiguousarray(s.mps, np.int32), np.ascontiguousarray(s.jxyz, np.int32),
                    np.ascontiguousarray(s.rxyz, np.float64), np.ascontiguousarray(s.ia_tbl, np.int32),
                    np.ascontiguousarray(s.rinv_uvu, np.float64), np.ascontiguousarray(s.uv, np.float64)]
        mps, jxyz, rxyz, ia, rinv, uv = self._nl
        _chk(load().salmon_hip_set_nonlocal(self.h, s.natom, s.nps, _p(mps, _ip), _p(jxyz, _ip), _p(rxyz), s.nlma,
                                            _p(ia, _ip), _p(rinv), _p(uv)), "set_nonlocal")

    def set_vlocal(self, V):
        v = np.ascontiguousarray(V, dtype=np.float64)
        if v.nbytes != self._gs:
            raise ValueError(f"set_vlocal: V_local(nx,ny,nz,nspin) of {self._gs // 8} values expected, got {v.size}")
        _chk(load().salmon_hip_set_vlocal(self.h, _p(v)), "set_vlocal")

    def set_vlocal_device(self, Vdev):
        _need(Vdev, self._gs, "set_vlocal_device")
        _chk(load().salmon_hip_set_vlocal_device(self.h, C.c_void_p(_ptr(Vdev))), "set_vlocal_device")

    def update_kvector(self, vec_ac):
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        _chk(load().salmon_hip_update_kvector(self.h, _p(a)), "update_kvector")

    def set_vec_ac(self, vec_ac):
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        _chk(load().salmon_hip_set_vec_ac(self.h, _p(a)), "set_vec_ac")

    def hpsi(self, tpsi, htpsi):
        _need(tpsi, self._orb, "hpsi tpsi", f64=False)
        _need(htpsi, self._orb, "hpsi htpsi", f64=False)
        _chk(load().salmon_hip_hpsi(self.h, C.c_void_p(_ptr(tpsi)), C.c_void_p(_ptr(htpsi))), "hpsi")

    def taylor(self, zc, tin, tout, work):
        for t, w in ((tin, "tin"), (tout, "tout"), (work, "work")):
            _need(t, self._orb, "taylor " + w, f64=False)
        z = np.ascontiguousarray(np.asarray(zc, dtype=np.complex128)).view(np.float64)
        _chk(load().salmon_hip_taylor(self.h, _p(z), len(z) // 2, C.c_void_p(_ptr(tin)), C.c_void_p(_ptr(tout)),
                                      C.c_void_p(_ptr(work))), "taylor")

    def taylor_reuse(self, zc, tin, tout, work, dv):
        """The Taylor step of tin when only V_local changed since the last taylor call on tin:
        H psi is updated by dv * psi (dv = V_now - V_then, device array) instead of recomputed."""
        for t, w in ((tin, "tin"), (tout, "tout"), (work, "work")):
            _need(t, self._orb, "taylor_reuse " + w, f64=False)
        _need(dv, self._gs, "taylor_reuse dv")
        z = np.ascontiguousarray(np.asarray(zc, dtype=np.complex128)).view(np.float64)
        _chk(load().salmon_hip_taylor_reuse(self.h, _p(z), len(z) // 2, C.c_void_p(_ptr(tin)), C.c_void_p(_ptr(tout)),
                                            C.c_void_p(_ptr(work)), C.c_void_p(_ptr(dv))), "taylor_reuse")

    def set_skip(self, flag):
        """Device-side gate: while set, hpsi/taylor, density and the potential update do nothing
        when the device int `flag` is non-zero (None clears)."""
        _chk(load().salmon_hip_set_skip(self.h, C.c_void_p(None if flag is None else _ptr(flag))), "set_skip")

    def copy_if(self, flag, dst, src):
        """dst = src (float64 device tensors) when the device int `flag` is non-zero."""
        n = dst.numel()
        assert src.numel() == n and dst.dtype == src.dtype
        _chk(load().salmon_hip_copy_if(self.h, C.c_void_p(_ptr(flag)), C.c_void_p(_ptr(dst)), C.c_void_p(_ptr(src)),
                                       n), "copy_if")

    def stream_copy(self, dst, src, reps=10):
        """Device ms per launch of the library's streaming copy of src into dst (torch tensors)."""
        nb = src.numel() * src.element_size()
        assert dst.numel() * dst.element_size() >= nb
        ms = C.c_double()
        _chk(load().salmon_hip_stream_copy(self.h, C.c_void_p(_ptr(dst)), C.c_void_p(_ptr(src)), nb, int(reps),
                                           C.byref(ms)), "stream_copy")
        return ms.value

    PREC = {"fp64": 0, "mixed": 1}

    def autotune(self, reps=3):
        """Pick the fastest stencil launch shape; returns (whole_plane, z_chunks, ms per launch)."""
        ch = C.c_int()
        ms = C.c_double()
        _chk(load().salmon_hip_autotune(self.h, int(reps), C.byref(ch), C.byref(ms)), "autotune")
        return ch.value // 10, ch.value % 10, ms.value

    def set_overlap(self, on=True):
        _chk(load().salmon_hip_set_overlap(self.h, int(bool(on))), "set_overlap")

    def _set_shape(self, whole_plane, z_chunks):
        _chk(load().salmon_hip_set_launch_shape(self.h, int(whole_plane), int(z_chunks)), "set_launch_shape")

    def set_precision(self, mode):
        """'fp64' (default) or 'mixed' (complex FP32 stencil chain, FP64 sums) for taylor."""
        _chk(load().salmon_hip_set_precision(self.h, self.PREC[mode]), "set_precision")

    def density(self, psi, rho):
        _need(psi, self._orb, "density psi", f64=False)
        _need(rho, self._gs, "density rho")
        _chk(load().salmon_hip_density(self.h, C.c_void_p(_ptr(psi)), C.c_void_p(_ptr(rho))), "density")

    def density_noalloc(self, psi, rho):
        """density with its partial sums in the library's per-GPU device array (no allocation)."""
        _need(psi, self._orb, "density_noalloc psi", f64=False)
        _need(rho, self._gs, "density_noalloc rho")
        _chk(load().salmon_hip_density_noalloc(self.h, C.c_void_p(_ptr(psi)), C.c_void_p(_ptr(rho))),
             "density_noalloc")

    def current(self, vec_ac, psi, curr):
        _need(psi, self._orb, "current psi", f64=False)
        _need(curr, 24 * self.s.nspin, "current curr")
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        _chk(load().salmon_hip_current(self.h, _p(a), C.c_void_p(_ptr(psi)), C.c_void_p(_ptr(curr))), "current")

    def current_next(self, vec_ac, psi, curr, vec_ac_next):
        """current of psi, and the phase table of vec_ac_next with psi's projection under it, built
        ahead for the next update_kvector(vec_ac_next) and taylor/hpsi of psi."""
        _need(psi, self._orb, "current_next psi", f64=False)
        _need(curr, 24 * self.s.nspin, "current_next curr")
        a = np.ascontiguousarray(vec_ac, dtype=np.float64)
        an = np.ascontiguousarray(vec_ac_next, dtype=np.float64)
        _chk(load().salmon_hip_current_next(self.h, _p(a), C.c_void_p(_ptr(psi)), C.c_void_p(_ptr(curr)), _p(an)),
             "current_next")

    def dipole(self, rho, out):
        _need(rho, self._gs, "dipole rho")
        _need(out, 32, "dipole out")
        _chk(load().salmon_hip_dipole(self.h, C.c_void_p(_ptr(rho)), C.c_void_p(_ptr(out))), "dipole")

    def set_poisson(self, al):
        a = np.ascontiguousarray(al, dtype=np.float64)
        _chk(load().salmon_hip_set_poisson(self.h, _p(a)), "set_poisson")

    def hartree(self, rho, vh):
        _need(rho, self._g, "hartree rho")
        _need(vh, self._g, "hartree vh")
        _chk(load().salmon_hip_hartree(self.h, C.c_void_p(_ptr(rho)), C.c_void_p(_ptr(vh))), "hartree")

    def xc_pz(self, rho, vxc, exc=None):
        _need(rho, self._gs, "xc_pz rho_s")
        _need(vxc, self._gs, "xc_pz vxc")
        _need(exc, self._g, "xc_pz exc")
        _chk(load().salmon_hip_xc_pz(self.h, C.c_void_p(_ptr(rho)), C.c_void_p(_ptr(vxc)),
                                     C.c_void_p(_ptr(exc) if exc is not None else None)), "xc_pz")

    def update_vlocal(self, vpsl, vh, vxc, vmix=None, vout=None):
        for t, w, n in ((vpsl, "vpsl", self._g), (vh, "vh", self._g), (vxc, "vxc", self._gs), (vmix, "vmix", self._gs),
                        (vout, "vout", self._gs)):
            _need(t, n, "update_vlocal " + w)
        opt = [C.c_void_p(_ptr(x) if x is not None else None) for x in (vmix, vout)]
        _chk(load().salmon_hip_update_vlocal(self.h, C.c_void_p(_ptr(vpsl)), C.c_void_p(_ptr(vh)),
                                             C.c_void_p(_ptr(vxc)), *opt), "update_vlocal")

    def vlocal_extrapolate(self, vold1, vold2, vout=None):
        """aetrs: V = 3 V - 3 vold1 + vold2 on the handle's potential, history shifted in place."""
        for t, w in ((vold1, "vold1"), (vold2, "vold2"), (vout, "vout")):
            _need(t, self._gs, "vlocal_extrapolate " + w)
        _chk(load().salmon_hip_vlocal_extrapolate(self.h, C.c_void_p(_ptr(vold1)), C.c_void_p(_ptr(vold2)),
                                                  C.c_void_p(_ptr(vout) if vout is not None else None)),
             "vlocal_extrapolate")

    def update_potential(self, rho, vpsl, vmix=None, vout=None):
        for t, w, n in ((rho, "rho_s", self._gs), (vpsl, "vpsl", self._g), (vmix, "vmix", self._gs),
                        (vout, "vout", self._gs)):
            _need(t, n, "update_potential " + w)
        opt = [C.c_void_p(_ptr(x) if x is not None else None) for x in (vmix, vout)]
        _chk(load().salmon_hip_update_potential(self.h, C.c_void_p(_ptr(rho)), C.c_void_p(_ptr(vpsl)), *opt),
             "update_potential")

    KERNELS = {"hpsi": 0, "hpsi_taylor1": 1, "hpsi_taylor": 2, "hpsi_taylor_last": 3, "proj": 4,
               "proj_current": 5, "nl_scatter": 6, "kb_fused": 7}

    def prof_start(self, max_launches=100000):
        _chk(load().salmon_hip_prof_start(self.h, int(max_launches)), "prof_start")

    def prof_read(self, kernel):
        """(total device ms, launch count) of one kernel kind since prof_start."""
        ms = C.c_double()
        n = C.c_int()
        _chk(load().salmon_hip_prof_stop(self.h, self.KERNELS[kernel], C.byref(ms), C.byref(n)), "prof_stop")
        return ms.value, n.value

    def prof_end(self):
        _chk(load().salmon_hip_prof_end(self.h), "prof_end")

    def synchronize(self):
        _chk(load().salmon_hip_synchronize(self.h), "synchronize")


class Batch:
    """Batched step over several Handles (salmon_hip_batch_*)."""

    def __init__(self, handles, stream=None):
        L = load()
        self.handles = list(handles)
        arr = (C.c_void_p * len(self.handles))(*[h.h.value for h in self.handles])
        raw = getattr(stream, "cuda_stream", stream)
        b = C.c_void_p()
        _chk(L.salmon_hip_batch_create(arr, len(self.handles), C.c_void_p(raw or None), C.byref(b)), "batch_create")
        self.b = b

    def close(self):
        if getattr(self, "b", None) is not None and self.b.value:
            load().salmon_hip_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_system(self, i, psi_a, psi_b, work, rho, ac, nt, curr_t, dip_t):
        _chk(load().salmon_hip_batch_set_system(self.b, i, *[C.c_void_p(_ptr(x)) for x in (psi_a, psi_b, work, rho, ac)],
                                                int(nt), C.c_void_p(_ptr(curr_t)), C.c_void_p(_ptr(dip_t))),
             "batch_set_system")

    def prepare(self, zc):
        z = np.ascontiguousarray(np.asarray(zc, dtype=np.complex128)).view(np.float64)
        _chk(load().salmon_hip_batch_prepare(self.b, _p(z), len(z) // 2), "batch_prepare")

    def autotune(self, reps=3):
        ch = C.c_int()
        ms = C.c_double()
        _chk(load().salmon_hip_batch_autotune(self.b, int(reps), C.byref(ch), C.byref(ms)), "batch_autotune")
        return ch.value, ms.value

    def set_time(self, itt):
        _chk(load().salmon_hip_batch_set_time(self.b, int(itt)), "batch_set_time")

    def step(self):
        _chk(load().salmon_hip_batch_step(self.b), "batch_step")

    def capture(self):
        _chk(load().salmon_hip_batch_capture(self.b), "batch_capture")

    def replay(self, npairs):
        _chk(load().salmon_hip_batch_replay(self.b, int(npairs)), "batch_replay")
