"""Pins the oracle's Hartree, PZ-LDA and self-consistent RT loop.

* Hartree: closed form 4 pi/|G|^2 cos(G.r) for cosine densities (including the
  Nyquist mode and odd grids), numpy's FFT solution of the periodic Poisson
  equation G^2 V(G) = 4 pi rho(G) with G = 0 removed, zero for constant rho;
* PZ81: exchange -0.458165/rs and the correlation values of the published
  parametrisation at rs = 0.5, 1, 2, 5; v_xc = d(rho e_xc)/d rho by central
  differences on both branches of the parametrisation;
* PZ81 and its spin-polarised form: bitwise against exc_cor_pz / exc_cor_pz_sp compiled
  from the reference sources (oracle/_ref, when built); for the polarised form also
  v_s = d(rho e_xc)/d rho_s by central differences in each channel (a wrong sign of the
  spin term or a wrong f'(zeta) fails it), zeta = 0 reducing to the unpolarised PZ81,
  zeta = 1 to the published polarised parameters with exchange 2^(1/3) e_x, and the
  channels' exchange exchanging the potentials exactly;
* SCF loop: every step rebuilt from dense-matrix Taylor polynomials, numpy
  densities and the two pinned functionals, with and without the
  predictor-corrector, and with the 'aetrs' propagator (two half steps, V extrapolated); electron number never grows; nspin = 2 with both channels
  holding the nspin = 1 orbitals follows the nspin = 1 run.
"""
import dataclasses

import numpy as np
import pytest

from tests import dense


def _grid(n, al):
    x = [np.arange(m) * al[d] / m for d, m in enumerate(n)]
    return np.meshgrid(x[2], x[1], x[0], indexing="ij")  # Z, Y, X with shape (nz, ny, nx)


@pytest.mark.parametrize("n", [(16, 18, 20), (15, 12, 9)])
def test_hartree_cosine_closed_form(oracle_lib, n):
    al = np.array([10.0, 11.0, 12.5])
    Z, Y, X = _grid(n, al)
    rho = np.full(Z.shape, 0.3)
    ref = np.zeros_like(rho)
    modes = [((2, -3, 1), 1.0, np.cos), ((0, 1, 0), 0.5, np.sin), ((n[0] // 2, 0, 1), 0.25, np.cos)]
    for m, amp, f in modes:
        G = 2 * np.pi * np.array(m) / al
        ph = G[0] * X + G[1] * Y + G[2] * Z
        rho = rho + amp * f(ph)
        ref = ref + amp * 4 * np.pi / G.dot(G) * f(ph)
    v = oracle_lib.hartree(rho, al)
    assert np.abs(v - ref).max() < 1e-12 * np.abs(ref).max()


def test_hartree_matches_fft_poisson(oracle_lib):
    rng = np.random.default_rng(3)
    n = (12, 10, 14)
    al = np.array([9.0, 8.0, 10.0])
    rho = rng.random((n[2], n[1], n[0]))
    v = oracle_lib.hartree(rho, al)
    g = [2 * np.pi * np.fft.fftfreq(n[d], d=al[d] / n[d]) for d in range(3)]
    G2 = g[2][:, None, None] ** 2 + g[1][None, :, None] ** 2 + g[0][None, None, :] ** 2
    G2[0, 0, 0] = 1.0
    rg = np.fft.fftn(rho)
    rg[0, 0, 0] = 0.0
    ref = np.fft.ifftn(4 * np.pi * rg / G2).real
    assert np.abs(v - ref).max() < 1e-12 * np.abs(ref).max()
    # a constant density carries no Hartree potential (G = 0 removed)
    assert np.abs(oracle_lib.hartree(np.full((n[2], n[1], n[0]), 0.7), al)).max() < 1e-13


def _rho_of_rs(rs):
    return 3.0 / (4.0 * np.pi * np.asarray(rs, dtype=float) ** 3)


def test_pz_published_values(oracle_lib):
    rs = np.array([0.5, 1.0, 2.0, 5.0])
    v, e = oracle_lib.xc_pz(_rho_of_rs(rs))
    ex = -0.4581652932831429 / rs  # -(3/4)(9/(4 pi^2))^(1/3)/rs
    # Perdew-Zunger 1981 correlation (Hartree): high density A ln rs + B + C rs ln rs + D rs,
    # low density gamma/(1 + beta1 sqrt(rs) + beta2 rs)
    ec_pub = np.array([-0.0760500, -0.0596000, -0.0450912, -0.0283390])
    assert np.abs(e - ex - ec_pub).max() < 2e-7


@pytest.mark.parametrize("rs", [0.3, 0.7, 0.95, 1.05, 1.5, 3.0, 8.0])
def test_pz_potential_is_energy_derivative(oracle_lib, rs):
    rho = float(_rho_of_rs(rs))
    h = 1e-5 * rho
    r = np.array([rho - h, rho + h, rho])
    v, e = oracle_lib.xc_pz(r)
    dfd = (r[1] * e[1] - r[0] * e[0]) / (2 * h)
    assert abs(v[2] - dfd) < 1e-8 * abs(v[2])


def test_pz_bitwise_vs_reference_build(oracle_lib):
    if not oracle_lib.ref_xc_available():
        pytest.skip("oracle/_ref/libxc_ref.so not built (needs the reference sources and flang)")
    rng = np.random.default_rng(21)
    rho = np.concatenate([10 ** rng.uniform(-9, 1.5, 4000), [0.0, 1e-14, 3 / (4 * np.pi), 0.2387, 0.2388]])
    v, e = oracle_lib.xc_pz(rho)
    vr, er = oracle_lib.ref_xc_pz(rho)
    assert np.array_equal(v, vr) and np.array_equal(e, er)
    ru = 10 ** rng.uniform(-9, 1.5, 4000)
    rd = 10 ** rng.uniform(-9, 1.5, 4000)
    rd[:500] = 0.0          # fully polarised
    ru[500:1000] = 0.0      # fully polarised the other way
    rd[1000:1500] = ru[1000:1500]  # unpolarised
    rs = np.stack([np.append(ru, [0.0, 3 / (4 * np.pi)]), np.append(rd, [0.0, 0.0])])
    v, e = oracle_lib.xc_pz_sp(rs)
    vr, er = oracle_lib.ref_xc_pz_sp(rs)
    assert np.all(np.isfinite(vr))
    assert np.array_equal(v, vr) and np.array_equal(e, er)


@pytest.mark.parametrize("rs", [0.4, 0.9, 1.3, 4.0])
@pytest.mark.parametrize("zeta", [0.0, 0.35, -0.6, 0.9])
def test_pz_sp_potential_is_energy_derivative(oracle_lib, rs, zeta):
    rho = float(_rho_of_rs(rs))
    ru, rd = 0.5 * rho * (1 + zeta), 0.5 * rho * (1 - zeta)
    h = 1e-5 * rho
    pts = [(ru, rd), (ru - h, rd), (ru + h, rd), (ru, rd - h), (ru, rd + h)]
    r = np.array(pts).T.copy()
    v, e = oracle_lib.xc_pz_sp(r)
    tot = r.sum(axis=0) + 1e-20
    f = tot * e
    dup = (f[2] - f[1]) / (2 * h)
    ddn = (f[4] - f[3]) / (2 * h)
    assert abs(v[0, 0] - dup) < 1e-8 * abs(v[0, 0]), (v[0, 0], dup)
    assert abs(v[1, 0] - ddn) < 1e-8 * abs(v[1, 0]), (v[1, 0], ddn)
    # the two channels' potentials differ by the spin term whenever zeta != 0
    assert (abs(v[0, 0] - v[1, 0]) > 1e-4) == (zeta != 0.0)


def test_pz_sp_limits(oracle_lib):
    rs = np.array([0.5, 1.0, 2.0, 5.0])
    rho = _rho_of_rs(rs)
    # zeta = 0: the unpolarised functional, up to the density offsets of the two forms (1e-10
    # and 1e-20): a relative change of at most 1e-10/rho in rs
    v0, e0 = oracle_lib.xc_pz(rho)
    v, e = oracle_lib.xc_pz_sp(np.stack([0.5 * rho, 0.5 * rho]))
    bound = 1e-10 / rho
    assert np.all(np.abs(e - e0) <= bound * np.abs(e0))
    assert np.all(np.abs(v[0] - v0) <= bound * np.abs(v0)) and np.array_equal(v[0], v[1])
    # zeta = 1: exchange 2^(1/3) e_x and the polarised correlation of PZ81 (gamma -0.0843,
    # beta1 1.3981, beta2 0.2611; A 0.01555, B -0.0269, C 0.0007, D -0.0048)
    v1, e1 = oracle_lib.xc_pz_sp(np.stack([rho, np.zeros_like(rho)]))
    ex = -0.4581652932831429 / rs * 2 ** (1 / 3)
    ec_p = np.where(rs >= 1.0 + 1e-12, -0.0843 / (1 + 1.3981 * np.sqrt(rs) + 0.2611 * rs),
                    0.01555 * np.log(rs) - 0.0269 + 0.0007 * rs * np.log(rs) - 0.0048 * rs)
    assert np.abs(e1 - (ex + ec_p)).max() < 1e-12
    # exchanging the channels exchanges the potentials, bit for bit
    rng = np.random.default_rng(4)
    r = rng.random((2, 64)) * 0.3
    va, ea = oracle_lib.xc_pz_sp(r)
    vb, eb = oracle_lib.xc_pz_sp(r[::-1].copy())
    assert np.array_equal(va, vb[::-1]) and np.array_equal(ea, eb)


def test_scf_spin_channels_follow_unpolarised(oracle_lib):
    """nspin = 2 with both channels holding the nspin = 1 orbitals (one electron each): the
    channel densities are half the nspin = 1 density, the Hartree potential sees their sum and
    the polarised PZ at zeta = 0 equals the unpolarised one, so the predictor-corrector runs
    agree to the scale of the PZ density offsets (1e-10 against 1e-20: a relative change of
    1e-10/rho in the potential, ~1e-8 at this system's lowest densities)."""
    from salmon2_amd import system

    kw = dict(n=12, no=3, nkgrid=(1, 1, 1), seed=31, nt=4)
    s1 = system.make_si_system(nspin=1, **kw)
    s2 = system.make_si_system(nspin=2, **kw)
    s2 = dataclasses.replace(s2, psi=np.repeat(s1.psi, 2, axis=2))
    p1, c1, d1, r1, v1 = oracle_lib.Oracle(s1).rt_run_scf(True, nt=4)
    p2, c2, d2, r2, v2 = oracle_lib.Oracle(s2).rt_run_scf(True, nt=4)
    assert v2.shape == (2,) + v1.shape
    tol = 2e-8
    assert np.abs(v2[0] - v1).max() < tol * np.abs(v1).max() and np.array_equal(v2[0], v2[1])
    assert np.abs(c2 - c1).max() < tol * np.abs(c1).max()
    assert np.abs(d2 - d1).max() < tol * np.abs(d1).max()
    assert np.abs(p2[:, :, 0] - p1[:, :, 0]).max() < tol * np.abs(p1).max()
    assert np.array_equal(p2[:, :, 0], p2[:, :, 1])


def _dense_density(s, psi):
    rho = np.zeros((s.nz, s.ny, s.nx))
    for ik in range(s.nk):
        for io in range(s.no):
            rho += np.abs(psi[ik, io, 0][:, : s.ny, :]) ** 2 * s.rocc[0, ik, io] * s.wtk[ik]
    return rho


def _dense_step(s, H0, V, psi, dt, n_hamil):
    out = np.zeros_like(psi)
    for ik in range(s.nk):
        H = H0[ik] + np.diag(V.ravel())
        for io in range(s.no):
            v = dense.interior(psi[ik, io, 0], s)
            acc = v.copy()
            term = v.copy()
            for n in range(1, n_hamil + 1):
                term = (-1j * dt / n) * (H @ term)
                acc = acc + term
            out[ik, io, 0][:, : s.ny, :] = acc.reshape(s.nz, s.ny, s.nx)
    return out


@pytest.mark.parametrize("pc", [False, True])
def test_scf_loop_dense(oracle_lib, tiny, pc):
    s = tiny
    nt = 3
    O = oracle_lib.Oracle(s)
    p_or, cur, dip, rne, vloc = O.rt_run_scf(pc, nt=nt)

    vpsl = s.vlocal[0]
    ac = s.pulse.table(s.dt, nt)
    s0 = dataclasses.replace(s, vlocal=np.zeros_like(s.vlocal))

    def potential(rho):
        return vpsl + oracle_lib.hartree(rho, s.al) + oracle_lib.xc_pz(rho)[0]

    psi = np.array(s.psi)
    V = potential(_dense_density(s, psi))
    for itt in range(1, nt + 1):
        A = 0.5 * (ac[itt] + ac[itt - 1])
        H0 = [dense.full_hamiltonian(s0, s.vec_k[ik] + A) for ik in range(s.nk)]
        if pc:
            pred = _dense_step(s, H0, V, psi, s.dt, s.n_hamil)
            V = 0.5 * (V + potential(_dense_density(s, pred)))
        psi = _dense_step(s, H0, V, psi, s.dt, s.n_hamil)
        rho = _dense_density(s, psi)
        V = potential(rho)
        assert abs(rne[itt - 1] - rho.sum() * s.hvol) < 1e-12 * rne[itt - 1]
    scale = np.abs(psi).max()
    assert np.abs(p_or - psi).max() < 1e-12 * scale
    assert np.abs(vloc - V).max() < 1e-12 * np.abs(V).max()
    # the truncated Taylor series contracts the high-frequency part of random orbitals
    # (|1 - ix - x^2/2 + i x^3/6 + x^4/24| < 1 for 0 < x < 2 sqrt 2), so rNe never grows
    assert np.all(np.diff(rne) <= 1e-12 * rne[0])


@pytest.mark.parametrize("scf", [False, True])
def test_aetrs_loop_dense(oracle_lib, tiny, scf):
    """The 'aetrs' propagator (time_evolution_step.f90:119-123, 425-475) rebuilt step by step from
    dense Taylor polynomials: a half step (dt/2) with V(t) and A = Ac_tot(itt), V extrapolated
    3V - 3V(t-dt) + V(t-2dt) (history starting at the initial V), a half step with it and
    A = Ac(itt+1); the current with Ac_tot(itt)."""
    s = tiny
    nt = 3
    p_or, cur, dip, rne, ac_or, vloc = oracle_lib.Oracle(s).rt_run_full(scf=scf, nt=nt, propagator="aetrs")
    vpsl = s.vlocal[0]
    ac = s.pulse.table(s.dt, nt)
    s0 = dataclasses.replace(s, vlocal=np.zeros_like(s.vlocal))

    def potential(rho):
        return vpsl + oracle_lib.hartree(rho, s.al) + oracle_lib.xc_pz(rho)[0]

    def ham(A):
        return [dense.full_hamiltonian(s0, s.vec_k[ik] + A) for ik in range(s.nk)]

    psi = np.array(s.psi)
    V = potential(_dense_density(s, psi)) if scf else vpsl
    v1 = v2 = V
    for itt in range(1, nt + 1):
        psi = _dense_step(s, ham(ac[itt]), V, psi, s.dt / 2, s.n_hamil)
        if scf:
            V, v1, v2 = 3.0 * V - 3.0 * v1 + v2, V, v1
        psi = _dense_step(s, ham(ac[itt + 1]), V, psi, s.dt / 2, s.n_hamil)
        rho = _dense_density(s, psi)
        if scf:
            V = potential(rho)
        assert abs(rne[itt - 1] - rho.sum() * s.hvol) < 1e-12 * rne[itt - 1]
    assert np.abs(p_or - psi).max() < 1e-12 * np.abs(psi).max()
    if scf:
        assert np.abs(vloc - V).max() < 1e-12 * np.abs(V).max()
    # a different propagator: the trajectories part at the scale of the field's step
    _, cur_mp, _, _, _, _ = oracle_lib.Oracle(s).rt_run_full(scf=scf, nt=nt)
    assert np.abs(cur - cur_mp).max() > 1e-9 * np.abs(cur).max()


def test_aetrs_fixed_field_free_is_two_half_steps(oracle_lib, tiny):
    """Without a field and with a fixed potential an aetrs step is two Taylor half steps of the same
    H: its distance from exp(-iH dt) is 2 (E dt/2)^5/5! against (E dt)^5/5! for one full Taylor
    step, 2^4 times smaller (a half step of the wrong length fails this)."""
    import scipy.linalg

    s = dataclasses.replace(tiny, pulse=None)
    O = oracle_lib.Oracle(s)
    p_ae, *_ = O.rt_run_full(nt=1, propagator="aetrs")
    p_mp, *_ = O.rt_run_full(nt=1)
    err_ae = err_mp = 0.0
    for ik in range(s.nk):
        H = dense.full_hamiltonian(s, s.vec_k[ik])
        U = scipy.linalg.expm(-1j * s.dt * H)
        for io in range(s.no):
            ex = U @ dense.interior(s.psi[ik, io, 0], s)
            err_ae = max(err_ae, np.abs(dense.interior(p_ae[ik, io, 0], s) - ex).max())
            err_mp = max(err_mp, np.abs(dense.interior(p_mp[ik, io, 0], s) - ex).max())
    assert 12 < err_mp / err_ae < 20, (err_ae, err_mp)


def test_scf_pc_differs_from_fixed_potential(oracle_lib, tiny):
    """The potential update changes the trajectory (the test above would pass on a no-op otherwise)."""
    O = oracle_lib.Oracle(tiny)
    _, c_fix, _, _ = O.rt_run(nt=3)
    _, c_scf, _, _, _ = O.rt_run_scf(True, nt=3)
    assert np.abs(c_fix - c_scf).max() > 1e-6 * np.abs(c_fix).max()
