"""SALMON's restart / checkpoint data (src/io/checkpoint_restart.f90) read into and written
from this package's layouts.

A restart directory (``data_for_restart/``, ``data_for_restart_rt/`` or a
``checkpoint_rt_NNNNNN/``) holds, in the single-file form SALMON writes without
``yn_self_checkpoint``:

  info.bin        Fortran sequential unformatted records: nk, no, iter, nprocs (int32),
                  if_real_orbital (4-byte logical)            (write_bin, :333-345)
  occupation.bin  one record rocc(no, nk, nspin) float64      (write_bin, :348-353)
  wfn.bin         stream access, zwf(nx, ny, nz, nspin, no, nk) complex128, or
                  rwf(...) float64 when if_real_orbital, unpadded, Fortran order
                  (write_wavefunction 'single' :607-631 and the MPI subarray view of
                  distributed_rw_wavefunction :2044-2090 write the same layout)
  Vh_stock.bin    two records Vh_stock1/2(nx, ny, nz) float64 (RT only, write_Vh_stock :897-967)
  rtdata.bin      RT with trans_longi /= 'tr' (write_rtdata :1925-1957): three sequential
                  records, (Ac_tot(:,itt), Ac_tot(:,itt+1)), (Ac_ind(:,itt), Ac_ind(:,itt+1)),
                  Ac_ind(:,0), float64

or, with method_wf_distributor = 'slice', ``k_KKKKKK_ob_BBBBBB/wfn_ob_OOOOOO.dat`` files of
the nspin blocks of one orbital (write_sliced :633-675).  Sequential records carry 4-byte
length markers before and after the data; a record longer than 2^31 - 1 bytes is a chain
of sub-records whose markers are negative for all but the last (the gfortran convention).

Orbitals are returned as ``psi[ik, io, ispin, iz, iy, ix]`` with the y extent padded to
``ly`` (the layout the device path takes); a file with more orbitals or k-points than asked
for is read for its first ``no`` / ``nk`` (read_wavefunction keeps the leading ones).
"""
from __future__ import annotations

import os
import struct

import numpy as np

_MAXREC = 2 ** 31 - 1


def read_records(path):
    """All records of a Fortran sequential unformatted file (4-byte markers)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        chunks = []
        while True:
            (n,) = struct.unpack_from("<i", data, pos)
            ln = abs(n)
            chunks.append(data[pos + 4 : pos + 4 + ln])
            (m,) = struct.unpack_from("<i", data, pos + 4 + ln)
            if abs(m) != ln:
                raise ValueError(f"{path}: record markers {n} / {m} disagree at offset {pos}")
            pos += 8 + ln
            if n >= 0:
                break
        out.append(b"".join(chunks))
    return out


def write_record(f, payload: bytes):
    """One sequential unformatted record (sub-records beyond 2^31 - 1 bytes)."""
    pos = 0
    n = len(payload)
    while True:
        ln = min(n - pos, _MAXREC)
        last = pos + ln >= n
        mark = ln if last else -ln
        f.write(struct.pack("<i", mark))
        f.write(payload[pos : pos + ln])
        f.write(struct.pack("<i", ln if last else -ln))
        pos += ln
        if last:
            break


def read_info(directory):
    rec = read_records(os.path.join(directory, "info.bin"))
    if len(rec) < 5:
        raise ValueError(f"{directory}/info.bin: {len(rec)} records, expected 5")
    nk, no, itt, nprocs = (struct.unpack("<i", r[:4])[0] for r in rec[:4])
    if_real = struct.unpack("<i", rec[4][:4])[0] != 0
    return dict(nk=nk, no=no, iter=itt, nprocs=nprocs, if_real_orbital=if_real)


def read_restart(directory, n, nspin=1, no=None, nk=None, ly=None):
    """Orbitals, occupations and the iteration count of a restart directory.

    n = (nx, ny, nz); no / nk default to the file's.  Returns dict(psi, rocc, iter, nprocs,
    if_real_orbital, mk, mo) with rocc[ispin, ik, io]."""
    nx, ny, nz = n
    info = read_info(directory)
    mk, mo = info["nk"], info["no"]
    no = mo if no is None else no
    nk = mk if nk is None else nk
    if no > mo or nk > mk:
        raise ValueError(f"{directory}: asked for {no} orbitals x {nk} k-points, the file has {mo} x {mk}")
    ly = ny if ly is None else ly
    (occ,) = read_records(os.path.join(directory, "occupation.bin"))[:1]
    rocc_f = np.frombuffer(occ, dtype="<f8")
    if rocc_f.size != mo * mk * nspin:
        raise ValueError(f"{directory}/occupation.bin: {rocc_f.size} values, expected {mo * mk * nspin}")
    rocc = rocc_f.reshape(nspin, mk, mo)[:, :nk, :no].copy()
    dt = np.dtype("<f8") if info["if_real_orbital"] else np.dtype("<c16")
    psi = np.zeros((nk, no, nspin, nz, ly, nx), dtype=np.complex128)
    single = os.path.join(directory, "wfn.bin")
    if os.path.exists(single):
        want = mk * mo * nspin * nx * ny * nz * dt.itemsize
        if os.path.getsize(single) != want:
            raise ValueError(f"{single}: {os.path.getsize(single)} bytes, expected {want}")
        w = np.memmap(single, dtype=dt, mode="r", shape=(mk, mo, nspin, nz, ny, nx))
        psi[:, :, :, :, :ny, :] = w[:nk, :no]
        del w
    else:
        _read_sliced(directory, psi, dt, n, nspin, mo)
    return dict(psi=psi, rocc=rocc, iter=info["iter"], nprocs=info["nprocs"],
                if_real_orbital=info["if_real_orbital"], mk=mk, mo=mo)


def _read_sliced(directory, psi, dt, n, nspin, mo):
    nx, ny, nz = n
    nk, no = psi.shape[:2]
    blocks = {}
    for name in os.listdir(directory):
        if name.startswith("k_") and "_ob_" in name:
            k, b = name[2:].split("_ob_")
            blocks.setdefault(int(k), []).append(int(b))
    for ik in range(1, nk + 1):
        starts = sorted(blocks.get(ik, []))
        if not starts:
            raise ValueError(f"{directory}: no wave-function slices of k-point {ik}")
        for io in range(1, no + 1):
            nb = max(b for b in starts if b <= io)
            path = os.path.join(directory, f"k_{ik:06d}_ob_{nb:06d}", f"wfn_ob_{io:06d}.dat")
            w = np.fromfile(path, dtype=dt)
            if w.size != nspin * nx * ny * nz:
                raise ValueError(f"{path}: {w.size} values, expected {nspin * nx * ny * nz}")
            psi[ik - 1, io - 1, :, :, :ny, :] = w.reshape(nspin, nz, ny, nx)


def read_vh_stock(directory, n):
    """Vh_stock1, Vh_stock2 [iz, iy, ix] of an RT checkpoint."""
    nx, ny, nz = n
    rec = read_records(os.path.join(directory, "Vh_stock.bin"))
    return tuple(np.frombuffer(r, dtype="<f8").reshape(nz, ny, nx).copy() for r in rec[:2])


def write_rtdata(directory, ac_tot, ac_ind, ac_ind0):
    """rtdata.bin: ac_tot = (Ac_tot(itt), Ac_tot(itt+1)), ac_ind likewise (2 x 3 each), ac_ind0 = Ac_ind(0)."""
    with open(os.path.join(directory, "rtdata.bin"), "wb") as f:
        write_record(f, np.ascontiguousarray(ac_tot, dtype="<f8").reshape(6).tobytes())
        write_record(f, np.ascontiguousarray(ac_ind, dtype="<f8").reshape(6).tobytes())
        write_record(f, np.ascontiguousarray(ac_ind0, dtype="<f8").reshape(3).tobytes())


def read_rtdata(directory):
    """(Ac_tot(itt), Ac_tot(itt+1)) (2, 3), (Ac_ind(itt), Ac_ind(itt+1)) (2, 3) and Ac_ind(0) (3,)."""
    rec = read_records(os.path.join(directory, "rtdata.bin"))
    if len(rec) < 3 or len(rec[0]) != 48 or len(rec[1]) != 48 or len(rec[2]) != 24:
        raise ValueError(f"{directory}/rtdata.bin: records {[len(r) for r in rec]}, expected 48, 48, 24 bytes")
    f = [np.frombuffer(r, dtype="<f8") for r in rec[:3]]
    return f[0].reshape(2, 3).copy(), f[1].reshape(2, 3).copy(), f[2].copy()


def write_checkpoint(directory, psi, ny, rocc, itt, nprocs=1, if_real_orbital=False, vh_stock=None,
                     sliced_block=0):
    """Write info.bin, occupation.bin, the orbitals (wfn.bin, or slices of sliced_block
    orbitals when sliced_block > 0) and, given (vh1, vh2), Vh_stock.bin."""
    os.makedirs(directory, exist_ok=True)
    psi = np.asarray(psi)
    nk, no, nspin, nz, _, nx = psi.shape
    with open(os.path.join(directory, "info.bin"), "wb") as f:
        for v in (nk, no, int(itt), int(nprocs), 1 if if_real_orbital else 0):
            write_record(f, struct.pack("<i", v))
    with open(os.path.join(directory, "occupation.bin"), "wb") as f:
        write_record(f, np.ascontiguousarray(rocc, dtype="<f8").reshape(nspin, nk, no).tobytes())
    body = psi[:, :, :, :, :ny, :]
    body = body.real.astype("<f8") if if_real_orbital else body.astype("<c16")
    if sliced_block <= 0:
        with open(os.path.join(directory, "wfn.bin"), "wb") as f:
            for ik in range(nk):
                for io in range(no):
                    f.write(np.ascontiguousarray(body[ik, io]).tobytes())
    else:
        nbo = min(no, sliced_block)
        for ik in range(nk):
            for io in range(no):
                nb = (io // nbo) * nbo + 1
                d = os.path.join(directory, f"k_{ik + 1:06d}_ob_{nb:06d}")
                os.makedirs(d, exist_ok=True)
                with open(os.path.join(d, f"wfn_ob_{io + 1:06d}.dat"), "wb") as f:
                    f.write(np.ascontiguousarray(body[ik, io]).tobytes())
    if vh_stock is not None:
        with open(os.path.join(directory, "Vh_stock.bin"), "wb") as f:
            for v in vh_stock:
                write_record(f, np.ascontiguousarray(v, dtype="<f8").tobytes())
