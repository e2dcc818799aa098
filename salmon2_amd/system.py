"""Synthetic periodic systems in SALMON's memory layout.

This replaces SALMON's ground-state restart plus pseudopotential set-up
(``src/atom/pp/prep_pp.f90``) for the seeded benchmark systems: diamond-cubic
Si supercells with a Gaussian local potential and Gaussian s+p
Kleinman-Bylander projectors.  Every array is returned in SALMON's Fortran
column-major order, expressed as a C-ordered numpy array with the axes
reversed, e.g. ``zwf(ix, iy, iz, ispin, io, ik)`` is ``psi[ik, io, ispin, iz, iy, ix]``.

Conventions reproduced from the reference:

* grid point ``i`` (1-based) of a periodic cell sits at ``(i - 1) * h``
  (``rshift = -h``, prep_pp.f90:125-129);
* ``hvol = det(A) / N_grid`` (lattice.f90:40);
* the orbital array has its y extent padded by one row
  (``USE_OPT_ARRAY_PADDING``, initialization.f90:606-609);
* projector spheres are enumerated per atom over periodic images
  ``j1, j2, j3`` in ``[-nc, nc]`` (outer to inner) and grid points
  ``z, y, x`` (outer to inner), keeping ``r < rps + 1e-12``
  (prep_pp.f90:80-106, 280-370), ``rxyz`` being the displacement from the
  image atom;
* ``uv = radial(r) * Ylm`` with SALMON's monomial ``Ylm``
  (salmon_math.f90:386-426: s = 1, p = (-y, z, -x) for m = -1, 0, 1);
* ``rinv_uvu = E * hvol`` (prep_pp.f90:518, ``inorm`` replaced by a seeded
  energy ``E``);
* Monkhorst-Pack k-points ``k = (i - 1/2)/n - 1/2`` in units of the
  reciprocal vectors, x fastest, ``wtk = 1/nk`` (lattice.f90:141-165);
* stencil coefficients of SALMON's ``Nd = 4`` (25-point) operator
  (initialization.f90:240-255, 1324-1371);
* Taylor coefficients ``zc(n) = (-i dt)^n / n!`` (initialization_rt.f90:492-508);
* the 'Asin2cos' pulse of em_field.f90:177-187 tabulated at ``t = i dt``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# cN(0:4, 4) and bN(1:4, 4): initialization.f90:1337-1340, 1367-1371
CN4 = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)
BN4 = (0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0)

# Diamond-cubic basis in reduced coordinates (samples/benchmark/bulk_Si/input_sc_Si.dat)
SI_BASIS = np.array(
    [
        [0.00, 0.00, 0.00],
        [0.25, 0.25, 0.25],
        [0.50, 0.00, 0.50],
        [0.00, 0.50, 0.50],
        [0.50, 0.50, 0.00],
        [0.75, 0.25, 0.75],
        [0.25, 0.75, 0.75],
        [0.75, 0.75, 0.25],
    ]
)
SI_A = 10.26  # bohr


@dataclasses.dataclass
class Pulse:
    """External vector potential A_ext(t) of calc_Ac_ext_t (src/rt/em_field.f90:86-306), A0 = f0/omega:

    'Asin2cos'      A = -epdir A0 sin^2(pi t/tw) cos(omega t + 2 pi phi)            (t < tw)
    'Acos2'..'Acos8' A = -A0 cos^n(pi tt/tw) Im(epdir exp(i(omega tt + 2 pi phi))),
                    tt = t - tw/2 - t1_start                                       (|tt| < tw/2)
    'impulse'       A = epdir e_impulse                                             (t >= 0)
    """

    A0: float
    omega: float
    tw: float
    phi_cep: float = 0.0
    epdir: tuple = (1.0, 0.0, 0.0)
    shape: str = "Asin2cos"
    t1_start: float = 0.0
    e_impulse: float = 1e-2

    def table(self, dt: float, nt: int) -> np.ndarray:
        """Ac_ext(1:3, 0:nt+1) as an (nt+2, 3) array (calc_Ac_ext_t(0, dt, 0, nt+1, ...))."""
        out = np.zeros((nt + 2, 3))
        acos = {"Acos2": 2, "Acos3": 3, "Acos4": 4, "Acos6": 6, "Acos8": 8}
        if self.shape not in ("Asin2cos", "impulse") and self.shape not in acos:
            raise ValueError(f"pulse shape {self.shape!r} not supported")
        for i in range(nt + 2):
            t = 0.0 + i * dt
            if self.shape == "Asin2cos":
                if t < self.tw:
                    s = math.sin(math.pi * t / self.tw)
                    env = s * s
                    c = math.cos(self.omega * t + self.phi_cep * 2.0 * math.pi)
                    for d in range(3):
                        out[i, d] = -self.epdir[d] * self.A0 * env * c
            elif self.shape == "impulse":
                if t >= 0.0:
                    for d in range(3):
                        out[i, d] = self.epdir[d] * self.e_impulse
            else:
                tt = t - 0.5 * self.tw - self.t1_start
                if abs(tt) < 0.5 * self.tw:
                    env = math.cos(math.pi * tt / self.tw) ** acos[self.shape]
                    sn = math.sin(self.omega * tt + self.phi_cep * 2.0 * math.pi)
                    for d in range(3):
                        out[i, d] = -self.A0 * env * (self.epdir[d] * sn)
        return out


@dataclasses.dataclass
class System:
    name: str
    nx: int
    ny: int
    nz: int
    ly: int
    al: np.ndarray  # (3,) cell lengths (orthogonal cell)
    hgs: np.ndarray  # (3,)
    hvol: float
    nspin: int
    no: int
    nk: int
    vec_k: np.ndarray  # (nk, 3)        == vec_k(3, nk)
    wtk: np.ndarray  # (nk,)
    rocc: np.ndarray  # (nspin, nk, no) == rocc(no, nk, nspin)
    coef_lap: np.ndarray  # (3, 4)      == coef_lap(4, 3)
    coef_nab: np.ndarray  # (3, 4)
    coef_lap0: float
    vlocal: np.ndarray  # (nspin, nz, ny, nx)
    rion: np.ndarray  # (natom, 3)
    natom: int
    nps: int
    mps: np.ndarray  # (natom,) int32
    jxyz: np.ndarray  # (natom, nps, 3) int32, 1-based
    rxyz: np.ndarray  # (natom, nps, 3)
    nlma: int
    ia_tbl: np.ndarray  # (nlma,) int32, 1-based
    rinv_uvu: np.ndarray  # (nlma,)
    uv: np.ndarray  # (nlma, nps)
    psi: np.ndarray  # (nk, no, nspin, nz, ly, nx) complex128
    dt: float = 0.08
    nt: int = 100
    n_hamil: int = 4
    pulse: Pulse | None = None
    seed: int = 0
    psi_rng_state: dict | None = None  # generator state before the orbital draws

    @property
    def ngrid(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def norb(self) -> int:
        return self.nk * self.no * self.nspin

    def ac_table(self) -> np.ndarray:
        if self.pulse is None:
            return np.zeros((self.nt + 2, 3))
        return self.pulse.table(self.dt, self.nt)

    def zc(self, aetrs: bool = False) -> np.ndarray:
        return taylor_coefs(self.dt, self.n_hamil, aetrs)


def stencil_coefs(hgs):
    """coef_lap(4,3), coef_nab(4,3), coef_lap0 of the orthogonal Nd=4 stencil."""
    hgs = [float(h) for h in hgs]
    lap0 = -0.5 * CN4[0] * (1.0 / hgs[0] ** 2 + 1.0 / hgs[1] ** 2 + 1.0 / hgs[2] ** 2)
    lap = np.array([[CN4[n + 1] / hgs[d] ** 2 for n in range(4)] for d in range(3)])
    nab = np.array([[BN4[n + 1] / hgs[d] for n in range(4)] for d in range(3)])
    return lap, nab, lap0


def taylor_coefs(dt: float, n_hamil: int, aetrs: bool = False) -> np.ndarray:
    """zc(n) = (-i dt)^n / n! (or dt/2 for aetrs) as complex128[n_hamil]."""
    h = dt / 2.0 if aetrs else dt
    out = np.zeros(n_hamil, dtype=np.complex128)
    for n in range(1, n_hamil + 1):
        z = complex(0.0, -h) ** n
        for m in range(1, n + 1):
            z = z / m
        out[n - 1] = z
    return out


def mp_kpoints(nkgrid, al):
    """Monkhorst-Pack grid, lattice.f90:141-165 (orthogonal cell)."""
    nk = int(np.prod(nkgrid))
    B = 2.0 * np.pi / np.asarray(al, dtype=float)
    k = np.zeros((nk, 3))
    for ik in range(nk):
        ix = ik % nkgrid[0] + 1
        iy = (ik // nkgrid[0]) % nkgrid[1] + 1
        iz = (ik // (nkgrid[0] * nkgrid[1])) % nkgrid[2] + 1
        red = [(ix - 0.5) / nkgrid[0] - 0.5, (iy - 0.5) / nkgrid[1] - 0.5, (iz - 0.5) / nkgrid[2] - 0.5]
        k[ik] = [red[d] * B[d] for d in range(3)]
    return k, np.full(nk, 1.0 / nk)


def si_positions(ncell):
    ncell = tuple(ncell)
    pos = []
    for cz in range(ncell[2]):
        for cy in range(ncell[1]):
            for cx in range(ncell[0]):
                for b in SI_BASIS:
                    pos.append([(cx + b[0]) * SI_A, (cy + b[1]) * SI_A, (cz + b[2]) * SI_A])
    return np.array(pos)


def gaussian_vlocal(n, hgs, al, rion, depth=-4.0, width=1.2):
    """V(r) = sum_atoms sum_{images in [-1,1]^3} depth * exp(-|r - R|^2 / width^2)."""
    nx, ny, nz = n
    x = np.arange(nx) * hgs[0]
    y = np.arange(ny) * hgs[1]
    z = np.arange(nz) * hgs[2]
    V = np.zeros((nz, ny, nx))
    w2 = width * width
    for R in rion:
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                for k in (-1, 0, 1):
                    dx = x - (R[0] + i * al[0])
                    dy = y - (R[1] + j * al[1])
                    dz = z - (R[2] + k * al[2])
                    ex = np.exp(-dx * dx / w2)
                    ey = np.exp(-dy * dy / w2)
                    ez = np.exp(-dz * dz / w2)
                    V += depth * (ez[:, None, None] * ey[None, :, None] * ex[None, None, :])
    return V


def _nc_images(rion, al, hgs, rps):
    """Image range nc(1:3) of prep_pp.f90:80-106 for a cuboid cell."""
    rps_max = rps + 1.5 * max(hgs) + 1e-2
    nc = []
    for d in range(3):
        rmin = rion[:, d].min()
        rmax = rion[:, d].max()
        if rmin + 2.0 * al[d] > al[d] + rps_max and rmax - 2.0 * al[d] < -rps_max:
            nc.append(1)
        else:
            nc.append(2)
    return nc, rps_max


def projector_spheres(n, hgs, al, rion, rps):
    """mps, jxyz (1-based) and rxyz per atom, in prep_pp.f90's enumeration order."""
    nx, ny, nz = n
    nc, rps_max = _nc_images(rion, al, hgs, rps)
    # grid coordinates u = i*h + rshift with rshift = -h  (i = 1..n)
    u = [np.array([i * hgs[d] + (-hgs[d]) for i in range(1, n[d] + 1)]) for d in range(3)]
    per_atom = []
    for R in rion:
        idx = []
        disp = []
        for j1 in range(-nc[0], nc[0] + 1):
            c1 = R[0] + j1 * al[0]
            if c1 <= 1 * hgs[0] - rps_max or c1 >= nx * hgs[0] + rps_max:
                continue
            for j2 in range(-nc[1], nc[1] + 1):
                c2 = R[1] + j2 * al[1]
                if c2 <= 1 * hgs[1] - rps_max or c2 >= ny * hgs[1] + rps_max:
                    continue
                for j3 in range(-nc[2], nc[2] + 1):
                    c3 = R[2] + j3 * al[2]
                    if c3 <= 1 * hgs[2] - rps_max or c3 >= nz * hgs[2] + rps_max:
                        continue
                    # tmpx = Rion + j1*al(1,1) + j2*al(1,2) + j3*al(1,3) (off-diagonals are 0)
                    tx = ((R[0] + j1 * al[0]) + j2 * 0.0) + j3 * 0.0
                    ty = ((R[1] + j1 * 0.0) + j2 * al[1]) + j3 * 0.0
                    tz = ((R[2] + j1 * 0.0) + j2 * 0.0) + j3 * al[2]
                    dx = u[0] - tx
                    dy = u[1] - ty
                    dz = u[2] - tz
                    r2 = (dx[None, None, :] * dx[None, None, :] + dy[None, :, None] * dy[None, :, None]) + dz[
                        :, None, None
                    ] * dz[:, None, None]
                    r = np.sqrt(r2)
                    iz, iy, ix = np.nonzero(r < rps + 1e-12)  # C order == z outer, x inner
                    if len(ix):
                        idx.append(np.stack([ix + 1, iy + 1, iz + 1], axis=1))
                        disp.append(np.stack([dx[ix], dy[iy], dz[iz]], axis=1))
        if idx:
            per_atom.append((np.concatenate(idx), np.concatenate(disp)))
        else:
            per_atom.append((np.zeros((0, 3), np.int64), np.zeros((0, 3))))
    natom = len(rion)
    mps = np.array([len(a[0]) for a in per_atom], dtype=np.int32)
    nps = int(mps.max()) if natom else 0
    jxyz = np.zeros((natom, nps, 3), dtype=np.int32)
    rxyz = np.zeros((natom, nps, 3))
    for ia, (ij, rr) in enumerate(per_atom):
        jxyz[ia, : len(ij)] = ij
        rxyz[ia, : len(ij)] = rr
    return mps, nps, jxyz, rxyz


def ylm_monomial(x, y, z, l, m):
    """SALMON's Ylm for l <= 1 (salmon_math.f90:386-426); for l = 2 and 3 real solid-harmonic
    polynomials without SALMON's normalisation (synthetic tables for the d- and f-channel paths:
    the GPU and the oracle consume the same tables, so only their shape matters)."""
    if l == 0:
        return np.ones_like(x)
    if l == 1:
        return {-1: -y, 0: z, 1: -x}[m]
    r2 = (x * x + y * y) + z * z
    if l == 2:
        return {-2: x * y, -1: y * z, 0: 3.0 * z * z - r2, 1: x * z, 2: x * x - y * y}[m]
    if l == 3:
        return {-3: y * (3.0 * x * x - y * y), -2: x * y * z, -1: y * (5.0 * z * z - r2),
                0: z * (5.0 * z * z - 3.0 * r2), 1: x * (5.0 * z * z - r2), 2: z * (x * x - y * y),
                3: x * (x * x - 3.0 * y * y)}[m]
    raise ValueError("channels up to l = 3 are generated")


def sp_projectors(mps, nps, rxyz, energies, hvol, rc=1.5, lmax=1):
    """uv(nps, nlma), ia_tbl, rinv_uvu for the channels l = 0 .. lmax per atom (set_lma / calc_uv
    order: atoms, then l, then m = -l .. l).  The s and p energies are the seeded ones; the d and
    f channels (lmax >= 2, synthetic) take smaller energies derived from them."""
    natom = len(mps)
    chans = [(l, mm) for l in range(lmax + 1) for mm in range(-l, l + 1)]
    nlma = len(chans) * natom
    uv = np.zeros((nlma, nps))
    ia_tbl = np.zeros(nlma, dtype=np.int32)
    rinv = np.zeros(nlma)
    lma = 0
    for ia in range(natom):
        m = int(mps[ia])
        x, y, z = rxyz[ia, :m, 0], rxyz[ia, :m, 1], rxyz[ia, :m, 2]
        r2 = (x * x + y * y) + z * z
        radial = np.exp(-r2 / (rc * rc))
        for l, mm in chans:
            uv[lma, :m] = radial * ylm_monomial(x, y, z, l, mm)
            ia_tbl[lma] = ia + 1
            # d and f: smaller strengths keep the KB spectrum inside the 4th-order Taylor step's
            # stable range at dt = 0.08 (their polynomials grow as r^l over the sphere)
            e = energies[ia, l] if l <= 1 else (0.5 * energies[ia, l % 2] + 0.25 * (l - 1)) * 0.3 ** (l - 1)
            rinv[lma] = e * hvol
            lma += 1
    return uv, ia_tbl, rinv, nlma


def random_orbitals(rng, n, ly, no, nk, nspin, hvol):
    """Complex-normal orbitals, Gram-Schmidt orthonormalised with the hvol inner product.

    The Gram-Schmidt result is obtained from a Householder QR whose column
    phases are fixed so that R has a positive real diagonal (the unique
    Gram-Schmidt factorisation)."""
    nx, ny, nz = n
    N = nx * ny * nz
    psi = np.zeros((nk, no, nspin, nz, ly, nx), dtype=np.complex128)
    sq = math.sqrt(hvol)
    for ik in range(nk):
        for isp in range(nspin):
            re = rng.standard_normal((no, N))
            im = rng.standard_normal((no, N))
            Z = (re + 1j * im).T  # (N, no)
            q, r = np.linalg.qr(Z * sq)
            d = np.diagonal(r)
            q = q * (d / np.abs(d))[None, :]
            orb = (q / sq).T.reshape(no, nz, ny, nx)
            psi[ik, :, isp, :, :ny, :] = orb
    return psi


def device_orbitals(s, device="cuda"):
    """The orbitals of ``random_orbitals`` for system s, drawn from the same generator state and
    orthonormalised on `device` (torch tensor, the padded layout of ``System.psi``).

    The Gram-Schmidt factor is formed as a Cholesky QR: G = Z^H Z hvol, R = chol(G)^H with a
    positive diagonal, Q = Z R^-1 -- the same unique factorisation as the host path's Householder
    QR with fixed column phases, to rounding (Z is a tall Gaussian matrix, condition ~1.1)."""
    import torch

    if s.psi_rng_state is None:
        raise ValueError("system has no recorded orbital generator state")
    rng = np.random.default_rng()
    rng.bit_generator.state = s.psi_rng_state
    nx, ny, nz = s.nx, s.ny, s.nz
    N = nx * ny * nz
    out = torch.zeros((s.nk, s.no, s.nspin, nz, s.ly, nx), dtype=torch.complex128, device=device)
    for ik in range(s.nk):
        for isp in range(s.nspin):
            re = rng.standard_normal((s.no, N))
            im = rng.standard_normal((s.no, N))
            Zt = torch.complex(torch.from_numpy(re), torch.from_numpy(im)).to(device)  # (no, N) = Z^T
            G = (Zt.conj() @ Zt.transpose(0, 1)).cpu().numpy() * s.hvol  # Z^H Z hvol
            L = np.linalg.cholesky(G)  # G = L L^H, R = L^H
            Rinv = np.linalg.inv(L.conj().T)
            # Q^T = (Z R^-1)^T = R^-T Z^T
            Qt = torch.from_numpy(np.ascontiguousarray(Rinv.T)).to(device) @ Zt
            out[ik, :, isp, :, :ny, :] = Qt.reshape(s.no, nz, ny, nx)
            del Zt, Qt
    return out


def make_si_system(
    ncell=(1, 1, 1),
    n=16,
    no=16,
    nkgrid=(1, 1, 1),
    seed=20261203,
    dt=0.08,
    nt=100,
    A0=0.01,
    omega=0.5,
    n_hamil=4,
    with_projectors=True,
    name="si",
    pad=1,
    rc=1.5,
    rcut=2.5,
    host_orbitals=True,
    nspin=1,
    lmax=1,
):
    """Seeded diamond-Si supercell of the kind the benchmark configurations name.

    host_orbitals=False leaves ``psi`` None and keeps the generator state in
    ``psi_rng_state``: ``device_orbitals`` then draws the same numbers and orthonormalises
    them on a GPU (the host QR of the 216-atom workload takes minutes)."""
    ncell = tuple(ncell)
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    al = np.array([SI_A * ncell[0], SI_A * ncell[1], SI_A * ncell[2]])
    hgs = al / np.array(nn, dtype=float)
    ngrid = nn[0] * nn[1] * nn[2]
    hvol = float(al[0] * al[1] * al[2]) / ngrid
    rion = si_positions(ncell)
    rng = np.random.default_rng(seed)
    lap, nab, lap0 = stencil_coefs(hgs)
    V = gaussian_vlocal(nn, hgs, al, rion)
    vec_k, wtk = mp_kpoints(nkgrid, al)
    nk = len(wtk)
    energies = rng.uniform(-2.0, 2.0, size=(len(rion), 2))
    if with_projectors:
        mps, nps, jxyz, rxyz = projector_spheres(nn, hgs, al, rion, rcut)
        uv, ia_tbl, rinv, nlma = sp_projectors(mps, nps, rxyz, energies, hvol, rc, lmax)
        natom = len(rion)
    else:
        natom, nps, nlma = 0, 0, 0
        mps = np.zeros(0, np.int32)
        jxyz = np.zeros((0, 0, 3), np.int32)
        rxyz = np.zeros((0, 0, 3))
        uv = np.zeros((0, 0))
        ia_tbl = np.zeros(0, np.int32)
        rinv = np.zeros(0)
    if nspin not in (1, 2):
        raise ValueError("nspin must be 1 or 2")
    ly = nn[1] + pad
    rng_state = rng.bit_generator.state
    psi = random_orbitals(rng, nn, ly, no, nk, nspin, hvol) if host_orbitals else None
    # spin-polarised (collinear): one electron per orbital and spin channel, and a spin-dependent
    # V_local (the second channel's differs by a smooth exchange-like splitting along x)
    rocc = np.full((nspin, nk, no), 2.0 / nspin)
    if nspin == 2:
        x = np.arange(nn[0]) * hgs[0]
        split = 0.1 * np.cos(2.0 * np.pi * x / al[0])[None, None, :]
        V = np.stack([V, V + split])
    pulse = Pulse(A0=A0, omega=omega, tw=nt * dt)
    return System(
        name=name,
        nx=nn[0],
        ny=nn[1],
        nz=nn[2],
        ly=ly,
        al=al,
        hgs=hgs,
        hvol=hvol,
        nspin=nspin,
        no=no,
        nk=nk,
        vec_k=vec_k,
        wtk=wtk,
        rocc=rocc,
        coef_lap=lap,
        coef_nab=nab,
        coef_lap0=lap0,
        vlocal=V[None].copy() if nspin == 1 else V.copy(),
        rion=rion,
        natom=natom,
        nps=nps,
        mps=mps,
        jxyz=jxyz,
        rxyz=rxyz,
        nlma=nlma,
        ia_tbl=ia_tbl,
        rinv_uvu=rinv,
        uv=uv,
        psi=psi,
        dt=dt,
        nt=nt,
        n_hamil=n_hamil,
        pulse=pulse,
        seed=seed,
        psi_rng_state=rng_state,
    )


def make_si_pp_system(pp_file, ncell=(1, 1, 1), n=12, no=16, nkgrid=(2, 2, 2), seed=20261212, dt=0.08, nt=100,
                      A0=0.01, omega=0.5, n_hamil=4, name="si_pp", pad=1, lmax_ps=2, lloc_ps=2):
    """Diamond-Si supercell with the tables of a real KY-format pseudopotential `pp_file`
    (e.g. SALMON's testsuites/pseudo/Si_rps.dat with Lmax_ps = Lloc_ps = 2, as its bulk-Si
    tests use):
    V_local = Vpsl and the KB projectors come from the projector-table builder
    (include/salmon_pp.h); the orbitals are seeded random as in make_si_system."""
    from . import pp as _pp

    ncell = tuple(ncell)
    nn = (n, n, n) if np.isscalar(n) else tuple(n)
    al = np.array([SI_A * ncell[0], SI_A * ncell[1], SI_A * ncell[2]])
    hgs = al / np.array(nn, dtype=float)
    hvol = float(al[0] * al[1] * al[2]) / (nn[0] * nn[1] * nn[2])
    rion = si_positions(ncell)
    kion = np.ones(len(rion), dtype=np.int32)
    P = _pp.PseudoPotential([pp_file], lloc_ps=[lloc_ps], lmax_ps=[lmax_ps])
    mps, nps, jxyz, rxyz, nlma, ia_tbl, rinv, uv = P.nonlocal_tables(kion, rion, al, nn)
    V = P.vpsl(kion, rion, al, nn)
    P.close()
    rng = np.random.default_rng(seed)
    lap, nab, lap0 = stencil_coefs(hgs)
    vec_k, wtk = mp_kpoints(nkgrid, al)
    nk = len(wtk)
    nspin = 1
    ly = nn[1] + pad
    psi = random_orbitals(rng, nn, ly, no, nk, nspin, hvol)
    return System(name=name, nx=nn[0], ny=nn[1], nz=nn[2], ly=ly, al=al, hgs=hgs, hvol=hvol, nspin=nspin, no=no,
                  nk=nk, vec_k=vec_k, wtk=wtk, rocc=np.full((nspin, nk, no), 2.0), coef_lap=lap, coef_nab=nab,
                  coef_lap0=lap0, vlocal=V[None].copy(), rion=rion, natom=len(rion), nps=nps, mps=mps,
                  jxyz=jxyz, rxyz=rxyz, nlma=nlma, ia_tbl=ia_tbl, rinv_uvu=rinv, uv=uv, psi=psi, dt=dt, nt=nt,
                  n_hamil=n_hamil, pulse=Pulse(A0=A0, omega=omega, tw=nt * dt), seed=seed)


# Named workloads (seed = 20261202 + configuration number)
def config(name: str, **over) -> System:
    """Named workloads.

    c1    : 8-atom Si, 16^3, 16 orbitals, Gamma, 100 steps, A0 = 0.01
    c2    : 64-atom Si (2x2x2), 32^3, 128 orbitals, 2x2x2 k, 1000 steps, A0 = 0.005
    c2k1  : c2 restricted to one k-point (parity tests of the full-size path)
    c4    : 216-atom Si (3x3x3), 48^3, 432 orbitals, 4 k (2x2x1), 500 steps
    tiny  : 8-atom Si, 12^3, 4 orbitals, 2 k, 10 steps (unit tests)
    """
    table = {
        "c1": dict(ncell=(1, 1, 1), n=16, no=16, nkgrid=(1, 1, 1), nt=100, A0=0.01, seed=20261203),
        "c2": dict(ncell=(2, 2, 2), n=32, no=128, nkgrid=(2, 2, 2), nt=1000, A0=0.005, seed=20261204),
        "c2k1": dict(ncell=(2, 2, 2), n=32, no=128, nkgrid=(1, 1, 1), nt=1000, A0=0.005, seed=20261204),
        "c4": dict(ncell=(3, 3, 3), n=48, no=432, nkgrid=(2, 2, 1), nt=500, A0=0.005, seed=20261206),
        "tiny": dict(ncell=(1, 1, 1), n=12, no=4, nkgrid=(1, 1, 2), nt=10, A0=0.01, seed=7),
    }
    kw = dict(table[name])
    kw.update(over)

    kw.setdefault("name", name)
    return make_si_system(**kw)


def batch_systems(n=24, seed=20261205, nt=100, host_orbitals=True):
    """Seeded uneven systems of the batched workload: Si supercells of 8..128 atoms, grid extents
    16..40 (step 4) chosen per axis so that the spacing stays in [0.55, 1.1] bohr (the 4th-order
    Taylor step with dt = 0.08 is stable for |dt E| < 2 sqrt 2, E_max ~ 9.75/h^2 for random
    orbitals), 16..256 orbitals, 1..27 k-points."""
    a = 10.26
    rng = np.random.default_rng(seed)
    cells = [(1, 1, 1), (1, 1, 2), (2, 2, 1), (2, 2, 2), (2, 2, 4)]
    kgrids = [(1, 1, 1), (2, 1, 1), (2, 2, 1), (2, 2, 2), (3, 3, 3)]
    out = []
    for i in range(n):
        cell = cells[rng.integers(len(cells))]
        grid = []
        for m in cell:
            cand = [g for g in range(16, 44, 4) if 0.55 <= a * m / g <= 1.1]
            grid.append(int(rng.choice(cand)))
        kg = kgrids[rng.integers(len(kgrids))]
        no = int(rng.choice([16, 32, 64, 128, 256]))
        out.append(make_si_system(ncell=cell, n=tuple(grid), no=no, nkgrid=kg, seed=seed + 1 + i, nt=nt,
                                  name=f"batch{i}", host_orbitals=host_orbitals))
    return out
