"""GPU parity of the self-consistent potential update and the predictor-corrector step.

Hartree (set_poisson + hartree), PZ-LDA (xc_pz), update_vlocal and the fused
update_potential against the oracle on seeded densities, then full RT runs with
the potential rebuilt every step (yn_fix_func = 'n'), with and without the
predictor-corrector, against the oracle's loop.  Tolerances as for the plain
path: <= 1e-12 relative on single-call results, <= 1e-9 on time series.
"""
import numpy as np
import pytest

from salmon2_amd import system

pytestmark = pytest.mark.gpu

CALL_TOL = 1e-12
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


CASES = {
    "tiny": lambda: system.config("tiny"),
    "c1": lambda: system.config("c1"),
    "si64_32cube": lambda: system.make_si_system(ncell=(2, 2, 2), n=32, no=4, nkgrid=(1, 1, 1), seed=11),
    "ragged_20x26x28": lambda: system.make_si_system(n=(20, 26, 28), no=2, nkgrid=(1, 1, 1), seed=13),
    "si_48x40x36_nopad": lambda: system.make_si_system(ncell=(1, 1, 1), n=(48, 40, 36), no=2, pad=0, seed=17),
}


@pytest.mark.parametrize("name", list(CASES))
def test_potential_update(torch_dev, oracle_lib, name):
    torch = torch_dev
    s = CASES[name]()
    print(f"[gpu] potential case={name} grid={s.nx}x{s.ny}x{s.nz}", flush=True)
    from salmon2_amd.hip import Handle

    rng = np.random.default_rng(5)
    shape = (s.nz, s.ny, s.nx)
    rho = 0.02 + rng.random(shape) * 0.1
    rho[0, 0, 0] = 1e-7  # low-density branch of PZ (rs > 1) and the high one both occur
    vpsl = s.vlocal[0]
    vmix = rng.standard_normal(shape)
    h = Handle(s, stream=torch.cuda.current_stream())
    h.set_poisson(s.al)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in
         dict(rho=rho, vpsl=vpsl, vmix=vmix).items()}
    vh = torch.zeros(shape, dtype=torch.float64, device="cuda")
    vxc = torch.zeros_like(vh)
    exc = torch.zeros_like(vh)
    v1 = torch.zeros_like(vh)
    v2 = torch.zeros_like(vh)
    v3 = torch.zeros_like(vh)
    h.hartree(d["rho"], vh)
    h.xc_pz(d["rho"], vxc, exc)
    h.update_vlocal(d["vpsl"], vh, vxc, vmix=d["vmix"], vout=v1)
    h.update_potential(d["rho"], d["vpsl"], vout=v2)
    h.update_potential(d["rho"], d["vpsl"], vmix=d["vmix"], vout=v3)
    torch.cuda.synchronize()
    vh_ref = oracle_lib.hartree(rho, s.al)
    vxc_ref, exc_ref = oracle_lib.xc_pz(rho)
    v_ref = vpsl + vh_ref + vxc_ref
    errs = dict(hartree=rel(vh.cpu().numpy(), vh_ref), vxc=rel(vxc.cpu().numpy(), vxc_ref),
                exc=rel(exc.cpu().numpy(), exc_ref), vlocal_mix=rel(v1.cpu().numpy(), 0.5 * (vmix + v_ref)),
                potential=rel(v2.cpu().numpy(), v_ref), potential_mix=rel(v3.cpu().numpy(), 0.5 * (vmix + v_ref)))
    for k, v in errs.items():
        print(f"[parity] {name} {k}: rel err {v:.3e}", flush=True)
        assert v <= CALL_TOL, k
    h.close()


SPIN_CASES = {
    "spin2_16x14x12": lambda: system.make_si_system(n=(16, 14, 12), no=3, nkgrid=(1, 1, 2), nspin=2, seed=19),
    "spin2_si8_16cube": lambda: system.make_si_system(ncell=(1, 1, 1), n=16, no=4, nkgrid=(1, 1, 1), nspin=2, seed=23),
}


@pytest.mark.parametrize("name", list(SPIN_CASES))
def test_potential_update_spin(torch_dev, oracle_lib, name):
    """nspin = 2: spin-polarised PZ on the channel densities, the Hartree potential of their sum,
    V_local(ispin) = Vpsl + Vh + Vxc(ispin) and the per-channel predictor-corrector average."""
    torch = torch_dev
    s = SPIN_CASES[name]()
    print(f"[gpu] potential spin case={name} grid={s.nx}x{s.ny}x{s.nz}", flush=True)
    from salmon2_amd.hip import Handle

    rng = np.random.default_rng(6)
    shape = (s.nz, s.ny, s.nx)
    rho_s = 0.01 + rng.random((2,) + shape) * 0.05
    rho_s[1, 0, 0, :4] = 0.0          # fully polarised points
    rho_s[0, 1, 0, :4] = 0.0
    rho_s[1, 2, 0, :4] = rho_s[0, 2, 0, :4]  # unpolarised points
    rho_s[:, 3, 0, 0] = 1e-8          # low-density branch (rs > 1)
    vpsl = s.vlocal[0]
    vmix = rng.standard_normal((2,) + shape)
    h = Handle(s, stream=torch.cuda.current_stream())
    h.set_poisson(s.al)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in
         dict(rho=rho_s, vpsl=vpsl, vmix=vmix).items()}
    vh = torch.zeros(shape, dtype=torch.float64, device="cuda")
    vxc = torch.zeros((2,) + shape, dtype=torch.float64, device="cuda")
    exc = torch.zeros_like(vh)
    v1, v2, v3 = (torch.zeros_like(vxc) for _ in range(3))
    h.hartree(d["rho"][0] + d["rho"][1], vh)
    h.xc_pz(d["rho"], vxc, exc)
    h.update_vlocal(d["vpsl"], vh, vxc, vmix=d["vmix"], vout=v1)
    h.update_potential(d["rho"], d["vpsl"], vout=v2)
    h.update_potential(d["rho"], d["vpsl"], vmix=d["vmix"], vout=v3)
    torch.cuda.synchronize()
    vh_ref = oracle_lib.hartree(rho_s[0] + rho_s[1], s.al)
    vxc_ref, exc_ref = oracle_lib.xc_pz_sp(rho_s)
    v_ref = vpsl[None] + vh_ref[None] + vxc_ref
    errs = dict(hartree=rel(vh.cpu().numpy(), vh_ref), vxc=rel(vxc.cpu().numpy(), vxc_ref),
                exc=rel(exc.cpu().numpy(), exc_ref), vlocal_mix=rel(v1.cpu().numpy(), 0.5 * (vmix + v_ref)),
                potential=rel(v2.cpu().numpy(), v_ref), potential_mix=rel(v3.cpu().numpy(), 0.5 * (vmix + v_ref)))
    for k, v in errs.items():
        print(f"[parity] {name} {k}: rel err {v:.3e}", flush=True)
        assert v <= CALL_TOL, k
    # the two channels' potentials differ where the channel densities do
    assert np.abs(vxc_ref[0] - vxc_ref[1]).max() > 1e-3
    h.close()


@pytest.mark.parametrize("name,pc,nt", [("tiny", False, 10), ("c1", True, 100), ("si64_32cube", True, 4),
                                        ("spin2_16x14x12", True, 20), ("spin2_si8_16cube", False, 10)])
def test_rt_scf_run(torch_dev, oracle_lib, name, pc, nt):
    torch = torch_dev
    s = (CASES.get(name) or SPIN_CASES[name])()
    print(f"[gpu] rt scf case={name} pc={pc} steps={nt}", flush=True)
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s, scf=True, pc=pc)
    res = rt.run(nt)
    vfin = rt.vcur.cpu().numpy()
    O = oracle_lib.Oracle(s)
    psi_ref, cur_ref, dip_ref, rne_ref, v_ref = O.rt_run_scf(pc, nt=nt)
    errs = dict(current=rel(res["curr"], cur_ref), dipole=rel(res["dip"], dip_ref), rNe=rel(res["rne"], rne_ref))
    for k, v in errs.items():
        print(f"[parity] {name}-scf{'-pc' if pc else ''}-{nt} {k}: rel err {v:.3e}", flush=True)
        assert v <= SERIES_TOL, k
    if s.nspin == 2:
        # per-channel currents (SYSname_rt.data's Jm_u / Jm_d columns) add up to the total
        assert res["curr_s"].shape == (nt, 2, 3) and np.array_equal(res["curr_s"].sum(axis=1), res["curr"])
        assert np.abs(res["curr_s"][:, 0] - res["curr_s"][:, 1]).max() > 1e-3 * np.abs(res["curr"]).max()
    e_psi = rel(res["psi"].cpu().numpy()[..., : s.ny, :], psi_ref[..., : s.ny, :])
    e_v = rel(vfin, v_ref)
    print(f"[parity] {name} final psi {e_psi:.3e} V_local {e_v:.3e}", flush=True)
    assert e_psi <= 1e-10 and e_v <= 1e-10


@pytest.mark.parametrize("name,prec", [("c1", "fp64"), ("si64_32cube", "fp64"), ("c1", "mixed")])
def test_taylor_reuse_equals_full_step(torch_dev, name, prec):
    """The corrector's shortcut: H psi updated by (V2 - V1) psi, then the remaining terms,
    equals a Taylor step computed from scratch with V2 (FP64: to round-off; mixed: within
    the FP32 chain's own rounding)."""
    torch = torch_dev
    s = CASES[name]()
    print(f"[gpu] taylor_reuse case={name} prec={prec}", flush=True)
    from salmon2_amd.hip import Handle

    rng = np.random.default_rng(8)
    v1 = s.vlocal[0]
    v2 = v1 + 0.05 * rng.standard_normal(v1.shape)
    h = Handle(s, stream=torch.cuda.current_stream())
    h.set_precision(prec)
    h.update_kvector(np.array([0.01, 0.0, -0.02]))
    psi = torch.from_numpy(s.psi).cuda()
    out1, out2, w1, w2 = (torch.zeros_like(psi) for _ in range(4))
    dv = torch.from_numpy(np.ascontiguousarray(v2 - v1)).cuda()
    h.set_vlocal(v1)
    h.taylor(s.zc(), psi, out1, w1)
    h.set_vlocal(v2)
    h.taylor_reuse(s.zc(), psi, out1, w1, dv)
    h.taylor(s.zc(), psi, out2, w2)
    torch.cuda.synchronize()
    e = rel(out1.cpu().numpy()[..., : s.ny, :], out2.cpu().numpy()[..., : s.ny, :])
    print(f"[parity] taylor_reuse {name} {prec}: rel err {e:.3e}", flush=True)
    assert e <= (1e-13 if prec == "fp64" else 2e-7)
    h.close()
