"""The device path on a system built from a real pseudopotential (SALMON's Si_rps.dat with
Lmax_ps = Lloc_ps = 2 through the projector-table builder): one hpsi, a fixed-potential RT
run and a self-consistent predictor-corrector run against the oracle on the same tables.
Tolerances as for the synthetic systems: 1e-12 on hpsi, 1e-9 on the time series."""
import numpy as np
import pytest

from salmon2_amd import system
from tests.conftest import SI_RPS

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


@pytest.fixture(scope="module")
def si_pp():
    # the bulk-Si testsuite geometry: 8 atoms, a = 10.26 bohr, 12^3 grid
    return system.make_si_pp_system(SI_RPS, n=12, no=8, nkgrid=(1, 1, 2), nt=20)


def test_hpsi_real_pp(torch_dev, oracle_lib, si_pp):
    torch = torch_dev
    s = si_pp
    print(f"[gpu] hpsi real Si pseudopotential, nlma={s.nlma} nps={s.nps}", flush=True)
    from salmon2_amd.hip import Handle

    A = np.array([0.01, -0.02, 0.015])
    h = Handle(s, stream=torch.cuda.current_stream())
    h.update_kvector(A)
    t = torch.from_numpy(s.psi).cuda()
    out = torch.zeros_like(t)
    h.hpsi(t, out)
    torch.cuda.synchronize()
    O = oracle_lib.Oracle(s)
    O.update_kvector(A)
    ref = O.hpsi(A, s.psi)
    e = rel(out.cpu().numpy()[..., : s.ny, :], ref[..., : s.ny, :])
    print(f"[parity] real-pp hpsi: rel err {e:.3e}", flush=True)
    assert e <= 1e-12
    h.close()


@pytest.mark.parametrize("scf", [False, True])
def test_rt_real_pp(torch_dev, oracle_lib, si_pp, scf):
    s = si_pp
    nt = 20 if not scf else 10
    print(f"[gpu] rt real Si pseudopotential scf={scf} steps={nt}", flush=True)
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s, scf=scf, pc=scf)
    res = rt.run(nt)
    O = oracle_lib.Oracle(s)
    if scf:
        _, cur, dip, rne, _ = O.rt_run_scf(True, nt=nt)
    else:
        _, cur, dip, rne = O.rt_run(nt=nt)
    errs = dict(current=rel(res["curr"], cur), dipole=rel(res["dip"], dip), rNe=rel(res["rne"], rne))
    for k, v in errs.items():
        print(f"[parity] real-pp {'scf-pc' if scf else 'fixed'} {k}: rel err {v:.3e}", flush=True)
        assert v <= 1e-9, k
