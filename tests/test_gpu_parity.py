"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerances (from the plain-FP64 contract): <= 1e-12 relative on hpsi and on
one Taylor step, <= 1e-9 relative on current / dipole / electron-count time
series after a run.  "relative" = max |gpu - ref| / max |ref| over the array.
"""
import numpy as np
import pytest

from salmon2_amd import system

pytestmark = pytest.mark.gpu

HPSI_TOL = 1e-12
SERIES_TOL = 1e-9


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


def _cases():
    return {
        "tiny": lambda: system.config("tiny"),
        "c1": lambda: system.config("c1"),
        "c1_noproj": lambda: system.config("c1", with_projectors=False, no=4),
        "si64_32cube": lambda: system.make_si_system(ncell=(2, 2, 2), n=32, no=12, nkgrid=(1, 1, 2), seed=11),
        "ragged_20x26x28": lambda: system.make_si_system(n=(20, 26, 28), no=5, nkgrid=(1, 1, 1), seed=13),
        "si_48x40x36_nopad": lambda: system.make_si_system(ncell=(1, 1, 1), n=(48, 40, 36), no=3, pad=0, seed=17),
        # 10-row y tiles with a partial last tile (rows 20-21), z chunks of 8 and 16 planes
        "ragged_24x22x16": lambda: system.make_si_system(n=(24, 22, 16), no=6, nkgrid=(1, 1, 1), seed=19),
        # projector spheres of radius 3.6: 58% of the touched points covered twice
        "overlap_rc3.6": lambda: system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), rcut=3.6, seed=23),
        # radius 5.3 > a/2: every touched point covered 2-10 times, also by images of the same atom
        "selfimage_rc5.3": lambda: system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), rcut=5.3, seed=23),
        # rows wider than 64 points: 3-row y tiles, halo ring of 4 chunks (8 x 72 halo points), two
        # halo elements per lane in the local current
        "wide_72x12x16": lambda: system.make_si_system(n=(72, 12, 16), no=3, nkgrid=(1, 1, 1), seed=29),
        # 136-point rows: one-row tiles, halo ring of 8 chunks, every y neighbour from the halo
        "wide_136x8x8": lambda: system.make_si_system(n=(136, 8, 8), no=2, nkgrid=(1, 1, 1), seed=31),
        # odd row length
        "odd_21x18x16": lambda: system.make_si_system(n=(21, 18, 16), no=4, nkgrid=(1, 1, 1), seed=37),
        # spin-polarised (collinear): two channels, each with its own V_local
        "spin2_16x14x12": lambda: system.make_si_system(n=(16, 14, 12), no=3, nkgrid=(1, 1, 2), seed=47, nspin=2),
        # d and f channels: 9 and 16 projectors per atom (the projection, fused KB pass and scatter
        # templates for more than 4 projectors per atom)
        "spd_12cube": lambda: system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), seed=53, lmax=2),
        "spdf_14cube": lambda: system.make_si_system(n=14, no=5, nkgrid=(1, 1, 2), seed=59, lmax=3),
    }


CASES = _cases()


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    return request.param, CASES[request.param]()


def _report(name, what, err):
    print(f"[parity] {name} {what}: rel err {err:.3e}", flush=True)


def test_hpsi(torch_dev, oracle_lib, case):
    torch = torch_dev
    name, s = case
    print(f"[gpu] hpsi case={name} grid={s.nx}x{s.ny}x{s.nz} ly={s.ly} norb={s.norb} natom={s.natom}", flush=True)
    from salmon2_amd.hip import Handle

    A = np.array([0.013, -0.021, 0.007])
    h = Handle(s, stream=torch.cuda.current_stream())
    h.update_kvector(A)
    t = torch.from_numpy(s.psi).cuda()
    out = torch.zeros_like(t)
    h.hpsi(t, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ref = O.hpsi(A, s.psi)
    err = rel(got[..., : s.ny, :], ref[..., : s.ny, :])
    _report(name, "hpsi", err)
    assert err <= HPSI_TOL
    if s.ly > s.ny:
        assert np.all(got[..., s.ny:, :] == 0)
    h.close()


def test_taylor(torch_dev, oracle_lib, case):
    torch = torch_dev
    name, s = case
    print(f"[gpu] taylor case={name}", flush=True)
    from salmon2_amd.hip import Handle

    A = np.array([-0.004, 0.009, 0.012])
    h = Handle(s, stream=torch.cuda.current_stream())
    h.update_kvector(A)
    tin = torch.from_numpy(s.psi).cuda()
    tout = torch.zeros_like(tin)
    work = torch.zeros_like(tin)
    h.taylor(s.zc(), tin, tout, work)
    torch.cuda.synchronize()
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ref_out, ref_in, ref_work = O.taylor(A, s.psi)
    got = tout.cpu().numpy()
    err = rel(got[..., : s.ny, :], ref_out[..., : s.ny, :])
    _report(name, "taylor", err)
    assert err <= HPSI_TOL
    # documented buffer contract: tspsi_in unchanged, sshtpsi = H tspsi_in
    assert np.array_equal(tin.cpu().numpy(), s.psi)
    H1 = O.hpsi(A, s.psi)
    err3 = rel(work.cpu().numpy()[..., : s.ny, :], H1[..., : s.ny, :])
    assert err3 <= HPSI_TOL
    h.close()


def test_observables(torch_dev, oracle_lib, case):
    torch = torch_dev
    name, s = case
    print(f"[gpu] observables case={name}", flush=True)
    from salmon2_amd.hip import Handle

    A_mid = np.array([0.02, 0.0, -0.01])
    A_now = np.array([0.025, 0.001, -0.012])
    h = Handle(s, stream=torch.cuda.current_stream())
    h.update_kvector(A_mid)
    psi = torch.from_numpy(s.psi).cuda()
    rho = torch.zeros((s.nspin, s.nz, s.ny, s.nx), dtype=torch.float64, device="cuda")
    cur = torch.zeros((s.nspin, 3), dtype=torch.float64, device="cuda")
    dip = torch.zeros(4, dtype=torch.float64, device="cuda")
    h.density(psi, rho)
    h.current(A_now, psi, cur)
    h.dipole(rho, dip)
    torch.cuda.synchronize()
    O = oracle_lib.Oracle(s)
    O.update_kvector(A_mid)
    rho_ref = O.density(s.psi)
    cur_ref = O.current(A_now, s.psi)
    dip_ref = O.dipole(rho_ref)
    e_rho = rel(rho.cpu().numpy(), rho_ref)
    e_cur = rel(cur.cpu().numpy(), cur_ref)
    e_dip = rel(dip.cpu().numpy(), dip_ref)
    _report(name, "density", e_rho)
    _report(name, "current", e_cur)
    _report(name, "dipole", e_dip)
    assert e_rho <= HPSI_TOL
    assert e_cur <= SERIES_TOL
    assert e_dip <= SERIES_TOL
    h.close()


RT_CASES = {
    # spin-polarised: two channels with their own V_local (J summed over the channels, dipole and
    # N_e of the total density)
    "spin2": lambda: system.make_si_system(n=(16, 14, 12), no=3, nkgrid=(1, 1, 2), seed=47, nspin=2, nt=30),
    # s+p+d+f projectors (16 per atom): the current's moments with the next step's projection
    "spdf": lambda: system.make_si_system(n=14, no=5, nkgrid=(1, 1, 2), seed=59, lmax=3, nt=30),
}


@pytest.mark.parametrize("name", list(RT_CASES))
def test_rt_fixed_potential_more_shapes(torch_dev, oracle_lib, name):
    """30 fixed-potential steps against the oracle: J(t), dipole(t), N_e(t), the final orbitals."""
    s = RT_CASES[name]()
    print(f"[gpu] rt {name} 30 steps", flush=True)
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s)
    res = rt.run()
    O = oracle_lib.Oracle(s)
    psi_ref, cur_ref, dip_ref, rne_ref = O.rt_run()
    e_cur = rel(res["curr"], cur_ref)
    e_dip = rel(res["dip"], dip_ref)
    e_rne = rel(res["rne"], rne_ref)
    e_psi = rel(res["psi"].cpu().numpy()[..., : s.ny, :], psi_ref[..., : s.ny, :])
    for k, v in (("current", e_cur), ("dipole", e_dip), ("rNe", e_rne), ("psi", e_psi)):
        _report(f"{name}-30", k, v)
    assert max(e_cur, e_dip, e_rne) <= SERIES_TOL
    assert e_psi <= HPSI_TOL
    rt.close()


def test_rt_c1_full_run(torch_dev, oracle_lib):
    """Configuration 1 (8-atom Si, 16^3, 16 orbitals, 100 steps): J(t), dipole(t), N_e(t)."""
    torch = torch_dev
    s = system.config("c1")
    print("[gpu] rt c1 100 steps", flush=True)
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s)
    res = rt.run()
    O = oracle_lib.Oracle(s)
    psi_ref, cur_ref, dip_ref, rne_ref = O.rt_run()
    e_cur = rel(res["curr"], cur_ref)
    e_dip = rel(res["dip"], dip_ref)
    e_rne = rel(res["rne"], rne_ref)
    e_psi = rel(res["psi"].cpu().numpy()[..., : s.ny, :], psi_ref[..., : s.ny, :])
    for k, v in (("current", e_cur), ("dipole", e_dip), ("rNe", e_rne), ("psi", e_psi)):
        _report("c1-100", k, v)
    assert e_cur <= SERIES_TOL and e_dip <= SERIES_TOL and e_rne <= SERIES_TOL
    assert e_psi <= 1e-10
    # the induced current is a real signal, not round-off
    assert np.abs(cur_ref).max() > 1e-6


@pytest.mark.parametrize("name", ["si64_32cube", "ragged_20x26x28", "ragged_24x22x16", "c1"])
def test_launch_shapes_keep_parity(torch_dev, oracle_lib, name):
    """Every stencil launch shape the autotuner may pick gives the same hpsi and Taylor step."""
    torch = torch_dev
    s = CASES[name]()
    print(f"[gpu] launch shapes case={name}", flush=True)
    from salmon2_amd.hip import Handle

    A = np.array([0.01, 0.02, -0.015])
    h = Handle(s, stream=torch.cuda.current_stream())
    h.update_kvector(A)
    plane, nzt, ms = h.autotune(reps=2)
    print(f"[gpu] autotune {name}: whole_plane={plane} z_chunks={nzt} {ms:.4f} ms", flush=True)
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ref_h = O.hpsi(A, s.psi)
    ref_t, _, _ = O.taylor(A, s.psi)
    t = torch.from_numpy(s.psi).cuda()
    for shape in [(1, 1), (1, 2), (0, 1), (0, 2), (0, 4), (-1, 0, "overlap")]:
        if shape[0] == 1 and s.nx * s.ny > 1024:
            continue
        h._set_shape(shape[0], shape[1])
        h.set_overlap(len(shape) > 2)
        out = torch.zeros_like(t)
        w = torch.zeros_like(t)
        h.hpsi(t, out)
        torch.cuda.synchronize()
        e1 = rel(out.cpu().numpy()[..., : s.ny, :], ref_h[..., : s.ny, :])
        h.taylor(s.zc(), t, out, w)
        torch.cuda.synchronize()
        e2 = rel(out.cpu().numpy()[..., : s.ny, :], ref_t[..., : s.ny, :])
        _report(name, f"shape {shape} hpsi/taylor", max(e1, e2))
        assert e1 <= HPSI_TOL and e2 <= HPSI_TOL
    h.close()


@pytest.mark.parametrize("n_hamil", [1, 2, 3, 6])
def test_taylor_orders(torch_dev, oracle_lib, n_hamil):
    """Taylor steps of order 1 .. 6 (chains of 0 .. 5 fused KB passes) against the oracle, FP64 and,
    from order 2, the mixed chain."""
    torch = torch_dev
    s = system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), rcut=3.6, seed=29, n_hamil=n_hamil)
    print(f"[gpu] taylor order {n_hamil}", flush=True)
    from salmon2_amd.hip import Handle

    A = np.array([0.012, -0.02, 0.007])
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ref, _, _ = O.taylor(A, s.psi)
    for prec, tol in (("fp64", HPSI_TOL), ("mixed", 1e-6)):
        if prec == "mixed" and n_hamil < 2:
            continue
        h = Handle(s, stream=torch.cuda.current_stream())
        h.set_precision(prec)
        h.update_kvector(A)
        t = torch.from_numpy(s.psi).cuda()
        out = torch.zeros_like(t)
        w = torch.zeros_like(t)
        h.taylor(s.zc(), t, out, w)
        torch.cuda.synchronize()
        e = rel(out.cpu().numpy()[..., : s.ny, :], ref[..., : s.ny, :])
        _report(f"order{n_hamil}", f"{prec} taylor", e)
        assert e <= tol
        h.close()
