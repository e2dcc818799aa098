"""Pins the oracle's density, current, dipole and pulse.

* density: sum rule sum(rho) hvol = sum rocc wtk and the direct |psi|^2 sum;
* current: plane-wave closed form of the kinetic current, zero current for
  real orbitals at Gamma with A = 0, the nonlocal current against dense
  projector matrices;
* dipole: closed form on a separable analytic density;
* A(t): the 'Asin2cos' closed form.
"""
import math

import numpy as np
import pytest

from salmon2_amd import system
from tests import dense


def test_density_sum_rule_and_direct(oracle_lib, tiny):
    s = tiny
    O = oracle_lib.Oracle(s)
    rho = O.density(s.psi)
    assert abs(rho.sum() * s.hvol - (s.rocc[0] * s.wtk[:, None]).sum()) < 1e-12
    ref = np.zeros_like(rho[0])
    for ik in range(s.nk):
        for io in range(s.no):
            ref += s.rocc[0, ik, io] * s.wtk[ik] * np.abs(s.psi[ik, io, 0, :, : s.ny]) ** 2
    assert np.abs(rho[0] - ref).max() / ref.max() < 1e-14


def _plane_wave_system(m, kvec, A0=0.0):
    s = system.make_si_system(n=12, no=1, nkgrid=(1, 1, 1), with_projectors=False, seed=3)
    s.vec_k = np.array([kvec], dtype=float)
    G = [2 * np.pi * m[d] / (s.al[d]) for d in range(3)]
    x = np.arange(s.nx) * s.hgs[0]
    y = np.arange(s.ny) * s.hgs[1]
    z = np.arange(s.nz) * s.hgs[2]
    ph = np.exp(1j * (G[0] * x[None, None, :] + G[1] * y[None, :, None] + G[2] * z[:, None, None]))
    ph /= np.sqrt(s.ngrid * s.hvol)
    s.psi = np.zeros_like(s.psi)
    s.psi[0, 0, 0, :, : s.ny] = ph
    return s, G


def test_current_plane_wave(oracle_lib):
    kvec = (0.1, -0.2, 0.05)
    s, G = _plane_wave_system((1, 2, -1), kvec)
    O = oracle_lib.Oracle(s)
    A = np.array([0.03, 0.0, -0.01])
    cur = O.current(A, s.psi)[0]
    norm = 1.0 / s.hvol  # sum |psi|^2 over the grid
    for d in range(3):
        h = s.hgs[d]
        grad = 2 * sum(system.BN4[n] * math.sin(n * G[d] * h) for n in range(1, 5)) / h
        ref = s.rocc[0, 0, 0] * s.wtk[0] * ((kvec[d] + A[d]) + grad) * norm / s.ngrid
        assert abs(cur[d] - ref) < 1e-12 * abs(norm / s.ngrid)
        # FD velocity tends to k + A + G
        assert abs(grad - G[d]) < 1e-2 * max(1.0, abs(G[d]))


def test_current_real_orbitals_gamma_zero(oracle_lib):
    s = system.make_si_system(n=12, no=3, nkgrid=(1, 1, 1), with_projectors=True, seed=5)
    s.psi = s.psi.real.astype(np.complex128)
    O = oracle_lib.Oracle(s)
    O.update_kvector([0.0, 0.0, 0.0])
    cur = O.current([0.0, 0.0, 0.0], s.psi)
    assert np.abs(cur).max() == 0.0


@pytest.mark.parametrize("lmax", [1, 2, 3])
def test_current_nonlocal_dense(oracle_lib, tiny, lmax):
    s = tiny if lmax == 1 else system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), seed=7, lmax=lmax)
    O = oracle_lib.Oracle(s)
    kA = np.array([0.05, -0.02, 0.01])
    O.update_kvector(kA)
    with_nl = O.current(kA, s.psi)[0]
    s0 = system.make_si_system(n=12, no=s.no, nkgrid=(1, 1, 2), with_projectors=False, seed=7)
    s0.psi = s.psi
    loc = oracle_lib.Oracle(s0).current(kA, s.psi)[0]
    jnl = np.zeros(3)
    for ik in range(s.nk):
        Z = dense.zekr_closed_form(s, s.vec_k[ik] + kA)
        P = dense.projector_matrix(s, Z)
        X = [np.zeros_like(Z) for _ in range(3)]
        for l in range(s.nlma):
            ia = s.ia_tbl[l] - 1
            m = s.mps[ia]
            for d in range(3):
                X[d][l, :m] = Z[l, :m] * s.rxyz[ia, :m, d]
        PX = [dense.projector_matrix(s, X[d]) for d in range(3)]
        for io in range(s.no):
            v = dense.interior(s.psi[ik, io, 0], s)
            u = (P.conj().T @ v) * s.rinv_uvu
            for d in range(3):
                ur = PX[d].conj().T @ v
                jnl[d] += s.rocc[0, ik, io] * s.wtk[ik] * 2 * np.sum(np.imag(np.conj(ur) * u))
    assert np.abs((with_nl - loc) - jnl / s.ngrid).max() < 1e-13 * np.abs(loc).max() + 1e-16


def test_dipole_closed_form(oracle_lib):
    s = system.make_si_system(n=(12, 10, 9), no=1, with_projectors=False, seed=1)
    O = oracle_lib.Oracle(s)
    fx = np.linspace(0.1, 1.0, s.nx)
    fy = np.linspace(1.0, 0.3, s.ny)
    fz = np.linspace(0.2, 0.9, s.nz)
    rho = (fz[:, None, None] * fy[None, :, None] * fx[None, None, :])[None]
    out = O.dipole(rho)
    pos = [np.arange(1, n + 1) - (0.5 if n % 2 == 0 else 0.0) for n in (s.nx, s.ny, s.nz)]
    tot = fx.sum() * fy.sum() * fz.sum()
    mom = [(pos[0] * fx).sum() * fy.sum() * fz.sum(), fx.sum() * (pos[1] * fy).sum() * fz.sum(),
           fx.sum() * fy.sum() * (pos[2] * fz).sum()]
    for d in range(3):
        assert abs(out[d] - (-mom[d] * s.hgs[d] * s.hvol)) < 1e-12 * abs(mom[d] * s.hgs[d] * s.hvol)
    assert abs(out[3] - tot * s.hvol) < 1e-13 * tot * s.hvol


def test_pulse_closed_form(oracle_lib):
    import ctypes as C

    p = system.Pulse(A0=0.01, omega=0.5, tw=8.0)
    dt, nt = 0.08, 100
    tab = p.table(dt, nt)
    ora = np.zeros((nt + 2, 3))
    ep = np.array([1.0, 0.0, 0.0])
    oracle_lib.lib().orc_ac_asin2cos(0.01, 0.5, 8.0, 0.0, ep.ctypes.data_as(C.POINTER(C.c_double)), dt, nt,
                                     ora.ctypes.data_as(C.POINTER(C.c_double)))
    t = np.arange(nt + 2) * dt
    ref = np.where(t < 8.0, -0.01 * np.sin(np.pi * t / 8.0) ** 2 * np.cos(0.5 * t), 0.0)
    assert np.abs(tab[:, 0] - ref).max() < 1e-17 and np.abs(ora - tab).max() == 0.0
    assert np.all(tab[:, 1:] == 0)
