"""Mixed-precision Taylor form against the FP64 oracle.

The complex-FP32 chain H psi, H^2 psi, ... carries FP32 rounding (relative
~6e-8 per term), so one step agrees with the FP64 oracle to ~1e-7 of |psi|;
over a run the error in the current, the dipole and the electron count is
stated against the oracle and bounded below.  The FP64 parts (projections,
KB term, Taylor sum, density, current) are exercised by the same calls.
"""
import numpy as np
import pytest

from salmon2_amd import system

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-6      # one Taylor step, relative to max |psi_out|
CUR_TOL = 2e-6       # J(t) after the 100-step c1 run, relative to max |J| (measured 6.5e-7)
DIP_TOL = 5e-8       # dipole(t), relative to its max (measured 7e-9)
NORM_TOL = 5e-8      # electron count, relative to the initial count (measured 7e-9)


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
    "si64_32cube": lambda: system.make_si_system(ncell=(2, 2, 2), n=32, no=12, nkgrid=(1, 1, 2), seed=11),
    "ragged_20x26x28": lambda: system.make_si_system(n=(20, 26, 28), no=5, nkgrid=(1, 1, 1), seed=13),
    "si_48x40x36_nopad": lambda: system.make_si_system(ncell=(1, 1, 1), n=(48, 40, 36), no=3, pad=0, seed=17),
    "ragged_24x22x16": lambda: system.make_si_system(n=(24, 22, 16), no=6, nkgrid=(1, 1, 1), seed=19),
    "overlap_rc3.6": lambda: system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), rcut=3.6, seed=23),
    "selfimage_rc5.3": lambda: system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), rcut=5.3, seed=23),
    # 72-point rows (separate-halo tiles, 4 halo chunks) and an odd row length (the middle terms'
    # 3-chunk slots with paired dword DMAs instead of the 2-chunk paired dwordx4 form)
    "wide_72x12x16": lambda: system.make_si_system(n=(72, 12, 16), no=3, nkgrid=(1, 1, 1), seed=29),
    "odd_21x18x16": lambda: system.make_si_system(n=(21, 18, 16), no=4, nkgrid=(1, 1, 1), seed=37),
    "spin2_16x14x12": lambda: system.make_si_system(n=(16, 14, 12), no=3, nkgrid=(1, 1, 2), seed=47, nspin=2),
    "spdf_14cube": lambda: system.make_si_system(n=14, no=5, nkgrid=(1, 1, 2), seed=59, lmax=3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_taylor_mixed_step(torch_dev, oracle_lib, name):
    torch = torch_dev
    s = CASES[name]()
    print(f"[gpu] mixed taylor case={name} grid={s.nx}x{s.ny}x{s.nz} norb={s.norb}", flush=True)
    from salmon2_amd.hip import Handle

    A = np.array([0.011, -0.017, 0.005])
    h = Handle(s, stream=torch.cuda.current_stream())
    h.set_precision("mixed")
    h.update_kvector(A)
    tin = torch.from_numpy(s.psi).cuda()
    keep = tin.clone()
    out = torch.zeros_like(tin)
    work = torch.zeros_like(tin)
    h.taylor(s.zc(), tin, out, work)
    torch.cuda.synchronize()
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ref, _, _ = O.taylor(A, s.psi)
    got = out.cpu().numpy()
    err = rel(got[..., : s.ny, :], ref[..., : s.ny, :])
    print(f"[parity] {name} mixed taylor step: rel err {err:.3e}", flush=True)
    assert err <= STEP_TOL
    assert err > 1e-12, "the FP32 chain is not in effect"
    assert torch.equal(tin, keep)
    if s.ly > s.ny:
        assert np.all(got[..., s.ny:, :] == 0)
    h.close()


def test_rt_mixed_c1(torch_dev, oracle_lib):
    s = system.config("c1")
    print("[gpu] rt mixed c1 100 steps", flush=True)
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s, precision="mixed")
    res = rt.run()
    O = oracle_lib.Oracle(s)
    psi_ref, cur_ref, dip_ref, rne_ref = O.rt_run()
    e_cur = rel(res["curr"], cur_ref)
    e_dip = rel(res["dip"], dip_ref)
    e_rne = float(np.abs(res["rne"] - rne_ref).max() / rne_ref[0])
    for k, v in (("current", e_cur), ("dipole", e_dip), ("rNe", e_rne)):
        print(f"[parity] c1-100 mixed {k}: rel err {v:.3e}", flush=True)
    assert e_cur <= CUR_TOL and e_dip <= DIP_TOL and e_rne <= NORM_TOL
