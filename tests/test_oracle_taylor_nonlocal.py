"""Pins the oracle's phase table, zpseudo and Taylor propagation.

* zekr_uV against its closed form conj(exp(i(k+A).r)) uv;
* zpseudo against a dense rank-nlma operator P diag(rinv) P^H (s+p, s+p+d and s+p+d+f channels);
* taylor against the truncated series sum_n (-i H dt)^n/n! psi of the dense
  Hamiltonian, against scipy's expm to O(dt^5), and its buffer contract.
"""
import numpy as np
import pytest
import scipy.linalg

from salmon2_amd import system
from tests import dense


def test_zekr_closed_form(oracle_lib, tiny):
    s = tiny
    O = oracle_lib.Oracle(s)
    kA = np.array([0.12, -0.3, 0.05])
    O.update_kvector(kA)
    for ik in range(s.nk):
        ref = dense.zekr_closed_form(s, s.vec_k[ik] + kA)
        assert np.abs(O.zekr[ik] - ref).max() < 1e-15


@pytest.mark.parametrize("lmax", [1, 2, 3])
def test_zpseudo_dense(oracle_lib, tiny, lmax):
    # s+p (the workloads'), and with d (9 projectors per atom) and f (16) channels
    s = tiny if lmax == 1 else system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), seed=7, lmax=lmax)
    assert s.nlma == (lmax + 1) ** 2 * s.natom
    O = oracle_lib.Oracle(s)
    kA = np.array([0.0, 0.07, -0.02])
    O.update_kvector(kA)
    h = O.zpseudo(s.psi, np.zeros_like(s.psi))
    for ik in range(s.nk):
        P = dense.projector_matrix(s, dense.zekr_closed_form(s, s.vec_k[ik] + kA))
        VNL = P @ np.diag(s.rinv_uvu) @ P.conj().T
        assert np.linalg.matrix_rank(VNL, tol=1e-10) <= s.nlma
        for io in range(s.no):
            ref = VNL @ dense.interior(s.psi[ik, io, 0], s)
            got = dense.interior(h[ik, io, 0], s)
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


def test_taylor_coefs(oracle_lib, tiny):
    import ctypes as C

    zc = np.zeros(8)
    oracle_lib.lib().orc_taylor_coefs(0.08, 4, 0, zc.ctypes.data_as(C.POINTER(C.c_double)))
    ref = np.array([(-1j * 0.08) ** n / np.math.factorial(n) if hasattr(np, "math") else 0 for n in range(1, 5)])
    import math

    ref = np.array([(-1j * 0.08) ** n / math.factorial(n) for n in range(1, 5)])
    assert np.abs(zc.view(np.complex128) - ref).max() < 1e-18
    assert np.abs(tiny.zc() - ref).max() < 1e-18


def test_taylor_dense_polynomial_and_expm(oracle_lib, tiny):
    s = tiny
    O = oracle_lib.Oracle(s)
    kA = np.array([0.02, -0.01, 0.03])
    O.update_kvector(kA)
    out, clob, work = O.taylor(kA, s.psi)
    dt = s.dt
    for ik in range(s.nk):
        H = dense.full_hamiltonian(s, s.vec_k[ik] + kA)
        U4 = np.eye(H.shape[0], dtype=np.complex128)
        term = np.eye(H.shape[0], dtype=np.complex128)
        for n in range(1, 5):
            term = term @ (-1j * dt * H) / n
            U4 = U4 + term
        Ue = scipy.linalg.expm(-1j * dt * H)
        emax = np.abs(np.linalg.eigvalsh(H)).max()
        for io in range(s.no):
            v = dense.interior(s.psi[ik, io, 0], s)
            got = dense.interior(out[ik, io, 0], s)
            assert np.abs(got - U4 @ v).max() / np.abs(v).max() < 1e-13
            # truncation error bounded by the first omitted term (emax dt)^5/5!
            assert np.linalg.norm(got - Ue @ v) <= (emax * dt) ** 5 / 120 * np.linalg.norm(v) * 1.01
        # buffer contract of taylor.f90: tspsi_in ends holding H^4 psi, sshtpsi H^3 psi
        H4v = np.linalg.matrix_power(H, 4) @ dense.interior(s.psi[ik, 0, 0], s)
        H3v = np.linalg.matrix_power(H, 3) @ dense.interior(s.psi[ik, 0, 0], s)
        assert np.abs(dense.interior(clob[ik, 0, 0], s) - H4v).max() / np.abs(H4v).max() < 1e-12
        assert np.abs(dense.interior(work[ik, 0, 0], s) - H3v).max() / np.abs(H3v).max() < 1e-12


def test_taylor_stable_at_dt(oracle_lib, c1):
    """4th-order Taylor is norm non-increasing iff |E dt| < 2*sqrt(2); check the spectral radius."""
    s = c1
    emax_kin = 3 * 0.5 * (-(np.array(__import__("salmon2_amd.system", fromlist=["CN4"]).CN4)[0]
                            + 2 * sum(((-1) ** n) * __import__("salmon2_amd.system", fromlist=["CN4"]).CN4[n]
                                      for n in range(1, 5)))) / s.hgs[0] ** 2
    assert emax_kin * s.dt < 2 * np.sqrt(2)
