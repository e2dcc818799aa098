"""Pins the oracle's stencil (zstencil_typical_seq / _omp) and hpsi.

* bitwise against SALMON's own zstencil_typical_{seq,omp}, compiled from the
  reference sources into oracle/_ref (skipped where the sources are absent);
* plane waves are exact eigenvectors with a closed-form eigenvalue;
* a constant orbital gives (V + |k+A|^2/2) psi since cN(0) + 2 sum cN(n) = 0;
* the whole hpsi equals a dense textbook Hamiltonian (tests/dense.py).
"""
import numpy as np
import pytest

from salmon2_amd import system
from tests import dense


def _kterms(s, kA):
    lap0 = s.coef_lap0 + 0.5 * float(np.dot(kA, kA))
    nabt = s.coef_nab * np.asarray(kA)[:, None]
    return lap0, nabt


def test_coefficients_match_reference_rule(oracle_lib):
    import ctypes as C

    hgs = np.array([0.64125, 0.7, 0.55])
    lap = np.zeros(12)
    nab = np.zeros(12)
    l0 = C.c_double()
    oracle_lib.lib().orc_stencil_coefs(hgs.ctypes.data_as(C.POINTER(C.c_double)),
                                       lap.ctypes.data_as(C.POINTER(C.c_double)),
                                       nab.ctypes.data_as(C.POINTER(C.c_double)), C.byref(l0))
    plap, pnab, pl0 = system.stencil_coefs(hgs)
    assert np.array_equal(lap, plap.ravel()) and np.array_equal(nab, pnab.ravel()) and l0.value == pl0
    # 8th-order central differences: sum rule and the second moment of the Laplacian
    cN = np.array(system.CN4)
    assert abs(cN[0] + 2 * cN[1:].sum()) < 1e-15
    assert abs(sum(cN[n] * n * n for n in range(1, 5)) - 1.0) < 1e-15
    bN = np.array(system.BN4)
    assert abs(2 * sum(bN[n] * n for n in range(1, 5)) - 1.0) < 1e-15


@pytest.mark.parametrize("mode", ["seq", "omp"])
@pytest.mark.parametrize("shape", [(12, 12, 12, 13), (16, 10, 8, 11), (5, 7, 9, 7)])
def test_stencil_bitwise_vs_reference(oracle_lib, mode, shape):
    if not oracle_lib.ref_available():
        pytest.skip("reference stencil not built (SALMON sources absent)")
    nx, ny, nz, ly = shape
    rng = np.random.default_rng(nx * 1000 + ny)
    psi = rng.standard_normal((nz, ly, nx)) + 1j * rng.standard_normal((nz, ly, nx))
    V = rng.standard_normal((nz, ny, nx))
    hgs = np.array([0.6, 0.7, 0.8])
    lap, nab, lap0 = system.stencil_coefs(hgs)
    kA = np.array([0.3, -0.1, 0.25])
    k_lap0 = lap0 + 0.5 * float(np.dot(kA, kA))
    nabt = nab * kA[:, None]
    h_orc = oracle_lib.zstencil(nx, ny, nz, ly, psi, V, k_lap0, lap, nabt, mode)
    h_ref = oracle_lib.ref_zstencil(nx, ny, nz, ly, psi, V, k_lap0, lap, nabt, mode)
    assert np.array_equal(h_orc[:, :ny, :], h_ref[:, :ny, :])


@pytest.mark.parametrize("mode", ["seq", "omp"])
def test_plane_wave_eigenvalue(oracle_lib, mode):
    nx, ny, nz, ly = 12, 10, 8, 11
    hgs = np.array([0.64, 0.7, 0.55])
    lap, nab, lap0 = system.stencil_coefs(hgs)
    m = (3, -2, 1)
    G = [2 * np.pi * m[d] / ((nx, ny, nz)[d] * hgs[d]) for d in range(3)]
    x = np.arange(nx) * hgs[0]
    y = np.arange(ny) * hgs[1]
    z = np.arange(nz) * hgs[2]
    ph = np.exp(1j * (G[0] * x[None, None, :] + G[1] * y[None, :, None] + G[2] * z[:, None, None]))
    psi = np.zeros((nz, ly, nx), dtype=np.complex128)
    psi[:, :ny, :] = ph
    V0 = -1.3
    V = np.full((nz, ny, nx), V0)
    kA = np.array([0.2, -0.15, 0.05])
    k_lap0, nabt = lap0 + 0.5 * float(kA @ kA), nab * kA[:, None]
    h = oracle_lib.zstencil(nx, ny, nz, ly, psi, V, k_lap0, lap, nabt, mode)
    cN, bN = system.CN4, system.BN4
    lam = V0 + 0.5 * float(kA @ kA)
    for d in range(3):
        hd = hgs[d]
        lam += -0.5 * (cN[0] + 2 * sum(cN[n] * np.cos(n * G[d] * hd) for n in range(1, 5))) / hd**2
        lam += 2 * kA[d] * sum(bN[n] * np.sin(n * G[d] * hd) for n in range(1, 5)) / hd
    err = np.abs(h[:, :ny] - lam * ph).max() / abs(lam)
    assert err < 1e-13
    # continuum limit: the FD symbol approaches |k + A + G|^2 / 2 for small G h
    assert abs(lam - V0 - 0.5 * sum((kA[d] + G[d]) ** 2 for d in range(3))) < 0.05


def test_constant_orbital(oracle_lib):
    nx, ny, nz, ly = 9, 8, 7, 9
    lap, nab, lap0 = system.stencil_coefs([0.5, 0.6, 0.7])
    psi = np.zeros((nz, ly, nx), dtype=np.complex128)
    psi[:, :ny] = 0.7 - 0.2j
    V = np.random.default_rng(0).standard_normal((nz, ny, nx))
    kA = np.array([0.1, 0.2, -0.3])
    h = oracle_lib.zstencil(nx, ny, nz, ly, psi, V, lap0 + 0.5 * float(kA @ kA), lap, nab * kA[:, None])
    expect = (V + 0.5 * float(kA @ kA)) * (0.7 - 0.2j)
    assert np.abs(h[:, :ny] - expect).max() < 1e-12


def test_padding_row_untouched(oracle_lib, tiny):
    O = oracle_lib.Oracle(tiny)
    O.update_kvector([0.01, 0.0, 0.0])
    h = O.hpsi([0.01, 0.0, 0.0], tiny.psi)
    assert np.all(h[..., tiny.ny:, :] == 0)


@pytest.mark.parametrize("kA", [(0.0, 0.0, 0.0), (0.31, -0.12, 0.07)])
def test_hpsi_equals_dense_hamiltonian(oracle_lib, tiny, kA):
    s = tiny
    O = oracle_lib.Oracle(s)
    O.update_kvector(kA)
    h = O.hpsi(kA, s.psi)
    for ik in range(s.nk):
        H = dense.full_hamiltonian(s, s.vec_k[ik] + np.asarray(kA))
        for io in range(s.no):
            v = dense.interior(s.psi[ik, io, 0], s)
            ref = H @ v
            got = dense.interior(h[ik, io, 0], s)
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


def test_hpsi_hermitian(oracle_lib, tiny):
    s = tiny
    O = oracle_lib.Oracle(s)
    kA = (0.05, 0.1, -0.2)
    O.update_kvector(kA)
    h = O.hpsi(kA, s.psi)
    a = dense.interior(s.psi[0, 0, 0], s)
    b = dense.interior(s.psi[0, 1, 0], s)
    ha = dense.interior(h[0, 0, 0], s)
    hb = dense.interior(h[0, 1, 0], s)
    assert abs(np.vdot(a, hb) - np.conj(np.vdot(b, ha))) < 1e-12 * np.abs(hb).max() * len(a)
