"""The learned-predictor step of the RT driver (salmon2_amd/rt.py with a Predictor), gated on
the device (salmon_unet_gate + salmon_hip_set_skip, no host synchronisation):

* gate closed (tol < 0): every attempt is rejected; each step is then the predictor-corrector
  step whose predictor is the attempt: its midpoint potential equals
  (V(t) + V[rho(Taylor(psi(t); V_pred))])/2 built by the oracle (density, Hartree, PZ) and
  its orbitals equal the oracle's Taylor step with that midpoint potential;
* gate open (tol = inf): every step after the four-step history is accepted, the run is
  reproducible bit for bit, and an accepted step equals one Taylor step of the oracle with
  the predicted potential;
* gate in between: accepted and rejected steps mix within one run, and every step's orbitals
  equal the oracle's Taylor step with the potential the step recorded as its midpoint.
Tolerance 1e-12 (the plain path's hpsi/Taylor bar); FP64 arithmetic."""
import dataclasses

import numpy as np
import pytest

from salmon2_amd import system

pytestmark = pytest.mark.gpu
NT = 10


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


def _predictor(torch, s, tol):
    from salmon2_amd.learned import Predictor

    P = Predictor(s.nx, s.ny, s.nz, seed=3, stream=torch.cuda.current_stream())
    P.tol_run = tol
    return P


def _run(torch, s, tol):
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s, scf=True, pc=True, precision="fp64", learned=_predictor(torch, s, tol), record=True)
    rt.setup_outputs(NT)
    snaps = []
    for _ in range(NT):
        snaps.append(rt.psi_a.clone())
        rt.step()
    return rt, rt.results(), snaps


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _step_vs_oracle(oracle_lib, s, rt, snaps, itt):
    """|psi(itt) - oracle Taylor(psi(itt-1); the step's recorded midpoint potential)|, relative."""
    vmid = rt.record[itt - 1][1].cpu().numpy()
    A = 0.5 * (rt.ac[itt] + rt.ac[itt - 1])
    O = oracle_lib.Oracle(dataclasses.replace(s, vlocal=vmid[None].copy()))
    O.update_kvector(A)
    ref, _, _ = O.taylor(A, snaps[itt - 1].cpu().numpy())
    got = (snaps[itt] if itt < len(snaps) else rt.psi_a).cpu().numpy()
    ny = s.ny
    return rel(got[..., :ny, :], ref[..., :ny, :])


def test_closed_gate_is_pc_with_the_attempt_as_predictor(torch_dev, oracle_lib):
    torch = torch_dev
    s = system.config("c1", nt=NT)
    print("[gpu] learned gate closed (device gate)", flush=True)
    rt, res, snaps = _run(torch, s, tol=-1.0)
    acc, att, _ = rt.predictor_stats()
    assert att == NT - 3 and acc == 0
    itt = NT
    vcur_prev = rt.record[itt - 2][0].cpu().numpy()
    vmid = rt.record[itt - 1][1].cpu().numpy()
    vpred = rt.record[itt - 1][2].cpu().numpy()
    A = 0.5 * (rt.ac[itt] + rt.ac[itt - 1])
    O = oracle_lib.Oracle(dataclasses.replace(s, vlocal=vpred[None].copy()))
    O.update_kvector(A)
    att_psi, _, _ = O.taylor(A, snaps[itt - 1].cpu().numpy())
    rho = O.density(att_psi)[0]
    vnew = s.vlocal[0] + oracle_lib.hartree(rho, s.al) + oracle_lib.xc_pz(rho)[0]
    e_v = rel(vmid, 0.5 * (vcur_prev + vnew))
    e_p = _step_vs_oracle(oracle_lib, s, rt, snaps, itt)
    print(f"[parity] learned closed gate: midpoint potential {e_v:.3e}, orbitals {e_p:.3e}", flush=True)
    assert e_v <= 1e-12 and e_p <= 1e-12


def test_open_gate_reproducible_and_accepted_step_matches_oracle(torch_dev, oracle_lib):
    torch = torch_dev
    s = system.config("c1", nt=NT)
    print("[gpu] learned gate open", flush=True)
    rt, res, snaps = _run(torch, s, tol=float("inf"))
    acc, att, _ = rt.predictor_stats()
    assert acc == att == NT - 3
    rt2, res2, _ = _run(torch, s, tol=float("inf"))
    for k in ("curr", "dip", "rne"):
        assert np.array_equal(res[k], res2[k]), k
    # accepted: the midpoint potential is the prediction itself
    assert np.array_equal(rt.record[NT - 1][1].cpu().numpy(), rt.record[NT - 1][2].cpu().numpy())
    e = _step_vs_oracle(oracle_lib, s, rt, snaps, NT)
    print(f"[parity] learned accepted step vs oracle Taylor: rel err {e:.3e}", flush=True)
    assert e <= 1e-12


def test_mixed_gate_every_step_matches_oracle(torch_dev, oracle_lib):
    torch = torch_dev
    s = system.config("c1", nt=NT)
    rt0, _, _ = _run(torch, s, tol=float("inf"))
    log = rt0.resid_log[: NT - 3].cpu().numpy()
    tol = float(np.median(log))
    print(f"[gpu] learned gate at tol {tol:.3e} (median of the open run's residuals)", flush=True)
    rt, _, snaps = _run(torch, s, tol=tol)
    acc, att, _ = rt.predictor_stats()
    print(f"[gpu] accepted {acc} of {att}", flush=True)
    assert att == NT - 3
    for itt in range(4, NT + 1):
        e = _step_vs_oracle(oracle_lib, s, rt, snaps, itt)
        assert e <= 1e-12, (itt, e)
