"""Oracle (test infrastructure only): CPU restatement of SALMON's pseudopotential set-up for
KY-format ('*_rps.dat') norm-conserving pseudopotentials on an orthogonal periodic cell.

Used by tests/ as the checker of the library's projector-table builder (salmon_pp_* in
include/salmon_pp.h); nothing in the product imports it.  Plain Python/numpy loops that
follow the reference routine by routine:

  read_ps_ky             src/atom/pp/input_pp.f90:288-341
  input_pp (post-read)   src/atom/pp/input_pp.f90:92-205   (radial grid, rps/nrps, anorm/inorm)
  making_ps_without_masking  src/atom/pp/input_pp.f90:1111-1248 (upp scaling, vloctbl, udvtbl)
  spline                 src/math/salmon_math.f90:494-584   (natural cubic spline)
  Ylm                    src/math/salmon_math.f90:386-426
  calc_nps / calc_jxyz   src/atom/pp/prep_pp.f90:195-370
  set_lma / calc_uv      src/atom/pp/prep_pp.f90:374-525
  calc_vpsl_periodic     src/atom/pp/prep_pp.f90:798-940  (the separable DFT branch)
  reciprocal grid        src/common/initialization.f90:676-700, 760-784

The derivative tables (dvpp, dupp, dudvtbl: forces only) and the masking branch are not
restated.  The spline's tridiagonal system is solved with numpy.linalg.solve where the
reference inverts it with LAPACK (DGETRF/DGETRI): the same equations.
"""
from __future__ import annotations

import math

import numpy as np

AU_LENGTH_AA = 0.52917721067  # src/io/inputoutput.f90:27
AU_ENERGY_EV = 27.21138505  # src/io/inputoutput.f90:26


def read_ps_ky(path, lmax_ps=-1, lloc_ps=0):
    """Element tables after read_ps_ky, input_pp's post-processing and
    making_ps_without_masking (dict of numpy arrays, radial index 1-based as in Fortran:
    index 0 of rad/vloctbl/udvtbl is unused)."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    mr, step, mlps, rzps = int(lines[0][0]), float(lines[0][1]), int(lines[0][2]), float(lines[0][3])
    zps = int(rzps + 1e-10)
    irpc = int(lines[1][0])
    del irpc  # partial-core radius: unused by the reference
    rrc = np.array([float(v) for v in lines[1][1 : 2 + mlps]])
    vpp = np.zeros((mr + 2, mlps + 1))
    upp = np.zeros((mr + 2, mlps + 1))
    for i in range(mr + 1):
        row = lines[2 + i]
        for l in range(mlps + 1):
            vpp[i, l] = float(row[2 + l])
    for i in range(mr + 1):
        row = lines[3 + mr + i]
        for l in range(mlps + 1):
            upp[i, l] = float(row[1 + l])
    # atomic units (input_pp.f90:320-324)
    step = step / AU_LENGTH_AA
    rrc = rrc / AU_LENGTH_AA
    vpp = vpp / AU_ENERGY_EV
    upp = upp * math.sqrt(AU_LENGTH_AA)
    nrmax = mr + 2
    rad = np.zeros(nrmax + 1)
    for i in range(1, nrmax + 1):
        rad[i] = (i - 1) * step
    if lmax_ps >= 0:
        mlps = lmax_ps
    lref = lloc_ps
    # input_pp.f90:114-119: rps = max rrc, nrps = first i with rad(i) > rps
    if np.any(rrc != 0.0):
        rps = float(np.max(rrc[: mlps + 1]))
    i = 1
    while i <= nrmax and not rad[i] > rps:
        i += 1
    nrps = i
    # anorm / inorm (input_pp.f90:163-178), before the radial functions are rescaled
    anorm = np.zeros(mlps + 1)
    inorm = np.zeros(mlps + 1, dtype=np.int64)
    for l in range(mlps + 1):
        a = 0.0
        for i in range(1, mr):
            r1 = rad[i + 1] - rad[i]
            a = a + (upp[i, l] ** 2 * (vpp[i, l] - vpp[i, lref]) + upp[i - 1, l] ** 2 * (vpp[i - 1, l] - vpp[i - 1, lref])) * r1
        a = 0.5 * a
        inorm[l] = 1
        if a < 0.0:
            a = -a
            inorm[l] = -1
        if abs(a) < 1e-10:
            inorm[l] = 0
        anorm[l] = math.sqrt(a)
    # making_ps_without_masking (input_pp.f90:1145-1151): u/r^(l+1) times sqrt((2l+1)/4pi)
    for l in range(mlps + 1):
        for i in range(1, mr + 1):
            upp[i, l] = upp[i, l] * math.sqrt((2 * l + 1.0) / (4 * math.pi)) / rad[i + 1] ** (l + 1)
        upp[0, l] = upp[1, l]
    vloctbl = np.zeros(nrps + 1)
    udvtbl = np.zeros((nrps + 1, mlps + 1))
    for l in range(mlps + 1):
        for i in range(1, nrps + 1):
            vloctbl[i] = vpp[i - 1, lref]
            udvtbl[i, l] = (vpp[i - 1, l] - vpp[i - 1, lref]) * upp[i - 1, l]
        if inorm[l] == 0:
            continue
        udvtbl[1 : nrps + 1, l] = udvtbl[1 : nrps + 1, l] / anorm[l]
    return dict(mr=mr, mlps=mlps, lref=lref, zps=zps, rps=rps, nrps=nrps, rad=rad, anorm=anorm, inorm=inorm,
                vloctbl=vloctbl, udvtbl=udvtbl)


def spline(xn, yn):
    """Natural cubic spline coefficients an, bn, cn, dn (0 .. Np-2) of salmon_math.f90:494-584."""
    Np = len(xn)
    Npm2 = Np - 2
    dxn = np.array([xn[i + 1] - xn[i] for i in range(Np - 1)])
    dyn = np.array([yn[i + 1] - yn[i] for i in range(Np - 1)])
    v = np.zeros(Npm2 + 1)  # 1-based
    for i in range(1, Npm2 + 1):
        v[i] = 6.0 * (dyn[i] / dxn[i] - dyn[i - 1] / dxn[i - 1])
    A = np.zeros((Npm2 + 1, Npm2 + 1))
    A[1, 1] = 2.0 * (dxn[1] + dxn[0])
    A[1, 2] = dxn[1]
    for i in range(2, Npm2):
        A[i, i + 1] = dxn[i]
        A[i, i] = 2.0 * (dxn[i] + dxn[i - 1])
        A[i, i - 1] = dxn[i - 1]
    A[Npm2, Npm2] = 2.0 * (dxn[Npm2] + dxn[Npm2 - 1])
    A[Npm2, Npm2 - 1] = dxn[Npm2 - 1]
    u = np.zeros(Npm2 + 2)
    u[1 : Npm2 + 1] = np.linalg.solve(A[1:, 1:], v[1:])
    bn = np.zeros(Npm2 + 1)
    bn[1 : Npm2 + 1] = 0.5 * u[1 : Npm2 + 1]
    an = np.zeros(Npm2 + 1)
    for i in range(Npm2):
        an[i] = (u[i + 1] - 2.0 * bn[i]) / (6.0 * dxn[i])
    an[Npm2] = (0.0 - 2.0 * bn[Npm2]) / (6.0 * dxn[Npm2])
    dn = np.array(yn[: Npm2 + 1], dtype=float)
    cn = np.zeros(Npm2 + 1)
    cn[0] = dyn[0] / dxn[0] - dxn[0] * (u[1] + 2.0 * 0.0) / 6.0
    for i in range(1, Npm2):
        cn[i] = dyn[i] / dxn[i] - dxn[i] * (u[i + 1] + 2.0 * u[i]) / 6.0
    cn[Npm2] = dyn[Npm2] / dxn[Npm2] - dxn[Npm2] * (0.0 + 2.0 * u[Npm2]) / 6.0
    return an, bn, cn, dn


def ylm(x, y, z, l, m):
    """SALMON's monomial Ylm (salmon_math.f90:386-426) for l <= 2."""
    r2 = x * x + y * y + z * z
    ilm = l * l + l + 1 + m
    s3 = math.sqrt(3.0)
    return {1: 1.0, 2: -y, 3: z, 4: -x, 5: s3 * x * y, 6: -s3 * y * z, 7: (3.0 * z * z - r2) / 2.0,
            8: -s3 * x * z, 9: math.sqrt(3.0 / 4.0) * (x * x - y * y)}[ilm]


def spheres(rps_elem, kion, rion, al, n):
    """calc_nps / calc_jxyz for an orthogonal periodic cell: per atom the grid points
    (1-based) and displacements within rps, images j1, j2, j3 outer, then z, y, x."""
    hgs = [al[d] / n[d] for d in range(3)]
    rps_max = max(rps_elem) + 1.5 * max(hgs) + 1e-2
    natom = len(kion)
    nc = [0, 0, 0]
    for d in range(3):
        rmin = min(rion[ia][d] for ia in range(natom))
        rmax = max(rion[ia][d] for ia in range(natom))
        nc[d] = 1 if (rmin + 2.0 * al[d] > al[d] + rps_max and rmax - 2.0 * al[d] < -rps_max) else 2
    out = []
    for ia in range(natom):
        rps = rps_elem[kion[ia]]
        pts = []
        for j1 in range(-nc[0], nc[0] + 1):
            if rion[ia][0] + j1 * al[0] <= 1 * hgs[0] - rps_max or rion[ia][0] + j1 * al[0] >= n[0] * hgs[0] + rps_max:
                continue
            for j2 in range(-nc[1], nc[1] + 1):
                if rion[ia][1] + j2 * al[1] <= 1 * hgs[1] - rps_max or rion[ia][1] + j2 * al[1] >= n[1] * hgs[1] + rps_max:
                    continue
                for j3 in range(-nc[2], nc[2] + 1):
                    if rion[ia][2] + j3 * al[2] <= 1 * hgs[2] - rps_max or rion[ia][2] + j3 * al[2] >= n[2] * hgs[2] + rps_max:
                        continue
                    tx = rion[ia][0] + j1 * al[0]
                    ty = rion[ia][1] + j2 * al[1]
                    tz = rion[ia][2] + j3 * al[2]
                    for i3 in range(1, n[2] + 1):
                        for i2 in range(1, n[1] + 1):
                            for i1 in range(1, n[0] + 1):
                                x = (i1 * hgs[0] - hgs[0]) - tx
                                y = (i2 * hgs[1] - hgs[1]) - ty
                                z = (i3 * hgs[2] - hgs[2]) - tz
                                r = math.sqrt(x * x + y * y + z * z)
                                if r < rps + 1e-12:
                                    pts.append((i1, i2, i3, x, y, z))
        out.append(pts)
    return out


def nonlocal_tables(elems, kion, rion, al, n):
    """set_lma + calc_uv: (mps, jxyz, rxyz, ia_tbl, rinv_uvu, uv) with jxyz/ia_tbl 1-based,
    uv[lma, j] (SALMON's uv(j, lma))."""
    hvol = al[0] * al[1] * al[2] / (n[0] * n[1] * n[2])
    sp = spheres([e["rps"] for e in elems], kion, rion, al, n)
    natom = len(kion)
    mps = np.array([len(p) for p in sp], dtype=np.int64)
    nps = int(mps.max()) if natom else 0
    jxyz = np.zeros((natom, nps, 3), dtype=np.int64)
    rxyz = np.zeros((natom, nps, 3))
    for ia, pts in enumerate(sp):
        for j, p in enumerate(pts):
            jxyz[ia, j] = p[:3]
            rxyz[ia, j] = p[3:]
    coefs = []
    for e in elems:
        nr = e["nrps"]
        xn = e["rad"][1 : nr + 1]
        cl = []
        for l in range(e["mlps"] + 1):
            cl.append(spline(xn, e["udvtbl"][1 : nr + 1, l]))
        coefs.append(cl)
    ia_tbl, rinv, rows = [], [], []
    for ia in range(natom):
        e = elems[kion[ia]]
        for ll in range(e["mlps"] + 1):
            if e["inorm"][ll] == 0:
                continue
            for m in range(-ll, ll + 1):
                ia_tbl.append(ia + 1)
                rinv.append(float(e["inorm"][ll]) * hvol)
                rows.append((ia, ll, m))
    uv = np.zeros((len(rows), nps))
    for lma, (ia, ll, m) in enumerate(rows):
        e = elems[kion[ia]]
        an, bn, cn, dn = coefs[kion[ia]][ll]
        radnl = e["rad"]
        for j in range(mps[ia]):
            x, y, z = rxyz[ia, j]
            r = math.sqrt(x * x + y * y + z * z) + 1e-50
            ir = 1
            while ir <= e["nrps"] and not radnl[ir] > r:
                ir += 1
            intr = ir - 1
            assert 0 <= intr < e["nrps"], "bad intr"
            xx = r - radnl[intr]
            s = intr - 1  # save_udvtbl_*(intr) = spline coefficient intr - 1
            uvr = an[s] * xx ** 3 + bn[s] * xx ** 2 + cn[s] * xx + dn[s]
            uv[lma, j] = uvr * ylm(x, y, z, ll, m)
    return mps, jxyz, rxyz, np.array(ia_tbl), np.array(rinv), uv


def vpsl_periodic(elems, kion, rion, al, n):
    """calc_vpsl_periodic: the local pseudopotential on the grid (separable DFT branch),
    returned as vpsl[iz, iy, ix]."""
    nx, ny, nz = n
    det = al[0] * al[1] * al[2]
    brl = [2.0 * math.pi / al[d] for d in range(3)]

    def gidx(i, N):  # initialization.f90:683-685
        return i - 1 - N * (1 + (1 if (i - 1 - (N + 1) // 2) >= 0 else -1)) // 2

    vg = np.zeros((nz, ny, nx), dtype=np.complex128)
    for iz in range(1, nz + 1):
        for iy in range(1, ny + 1):
            for ix in range(1, nx + 1):
                g = (gidx(ix, nx) * brl[0], gidx(iy, ny) * brl[1], gidx(iz, nz) * brl[2])
                g2 = g[0] ** 2 + g[1] ** 2 + g[2] ** 2
                gzero = (ix - 1) ** 2 + (iy - 1) ** 2 + (iz - 1) ** 2 == 0
                coef = 0.0 if gzero else 4.0 * math.pi / g2
                gs = math.sqrt(g2)
                zvg = []  # zVG_ion(G, element)
                for e in elems:
                    rad, vl, zps = e["rad"], e["vloctbl"], e["zps"]
                    s = 0.0
                    for i in range(2, e["nrps"] + 1):
                        r1 = 0.5 * (rad[i] + rad[i - 1])
                        dr = rad[i] - rad[i - 1]
                        va = 0.5 * (vl[i] + vl[i - 1])
                        if gzero:
                            s = s + 4.0 * math.pi * (r1 ** 2 * va + r1 * zps) * dr
                        else:
                            s = s + 4.0 * math.pi * math.sin(gs * r1) / gs * (r1 * va + zps) * dr
                    zvg.append(s)
                acc = 0j
                for ia in range(len(kion)):
                    zps = elems[kion[ia]]["zps"]
                    s = zvg[kion[ia]]
                    gd = g[0] * rion[ia][0] + g[1] * rion[ia][1] + g[2] * rion[ia][2]
                    acc = acc + (s - coef * zps) * (complex(math.cos(gd), -math.sin(gd)) / det)
                vg[iz - 1, iy - 1, ix - 1] = acc
    # inverse transform with egx(kx, ix) = exp(i 2 pi (ix-1)(kx-1) / N), z, y, then x
    def eg(N):
        k = np.arange(N)
        return np.exp(1j * (2.0 * math.pi * np.outer(k, k) / N))  # [k, i]

    ex, ey, ez = eg(nx), eg(ny), eg(nz)
    t = np.einsum("kyx,ki->iyx", vg, ez)
    t = np.einsum("zkx,ki->zix", t, ey)
    t = np.einsum("zyk,ki->zyi", t, ex)
    return t.real.copy()
