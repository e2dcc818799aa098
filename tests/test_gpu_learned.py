"""The learned-predictor step of the RT driver (salmon2_amd/rt.py with a Predictor):

* gate closed (tol = 0) with the plain fallback: every attempt is discarded and the run
  equals the plain predictor-corrector run bit for bit (the next step's inference runs on a
  side stream while the current kernels run; it must not disturb anything);
* gate open (tol = inf): every step after the four-step history is accepted, the run is
  reproducible bit for bit, and an accepted step equals one Taylor step of the oracle with
  the predicted potential (to the plain path's 1e-12);
* back-off: attempts only on the steps its schedule names, the others plain steps."""
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


def _run(torch, s, tol, fallback="plain", record=False, backoff=None):
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s, scf=True, pc=True, precision="fp64", learned=_predictor(torch, s, tol), fallback=fallback,
                      record=record, backoff=backoff)
    rt.setup_outputs(NT)
    snaps = []
    for _ in range(NT):
        snaps.append(rt.psi_a.clone())
        rt.step()
    return rt, rt.results(), snaps


def test_closed_gate_equals_plain_pc(torch_dev):
    torch = torch_dev
    from salmon2_amd.rt import RTPropagator

    s = system.config("c1", nt=NT)
    print("[gpu] learned gate closed vs plain predictor-corrector", flush=True)
    ref = RTPropagator(s, scf=True, pc=True).run(NT)
    rt, res, _ = _run(torch, s, tol=0.0)
    assert rt.attempted == NT - 3 and rt.accepted == 0
    for k in ("curr", "dip", "rne"):
        d = float(np.abs(res[k] - ref[k]).max())
        print(f"[parity] learned-closed {k}: max abs diff {d:.3e}", flush=True)
        assert d == 0.0, k


def test_backoff_attempts_and_equals_plain_pc(torch_dev):
    """Back-off (2 rejections, then every 3 steps): attempts at steps 4, 5 and 8 of 10; every
    step, attempted or not, is the plain predictor-corrector step."""
    torch = torch_dev
    from salmon2_amd.rt import RTPropagator

    s = system.config("c1", nt=NT)
    print("[gpu] learned back-off, gate closed", flush=True)
    ref = RTPropagator(s, scf=True, pc=True).run(NT)
    rt, res, _ = _run(torch, s, tol=0.0, backoff=(2, 3))
    assert rt.attempted == 3 and rt.accepted == 0 and rt.last_attempt == 8
    for k in ("curr", "dip", "rne"):
        assert np.array_equal(res[k], ref[k]), k


def test_open_gate_reproducible_and_accepted_step_matches_oracle(torch_dev, oracle_lib):
    torch = torch_dev
    s = system.config("c1", nt=NT)
    print("[gpu] learned gate open", flush=True)
    rt, res, snaps = _run(torch, s, tol=float("inf"), record=True)
    assert rt.accepted == rt.attempted == NT - 3
    _, res2, _ = _run(torch, s, tol=float("inf"))
    for k in ("curr", "dip", "rne"):
        assert np.array_equal(res[k], res2[k]), k
    # the last step was accepted: psi(t+dt) = Taylor(psi(t)) with V = the predicted midpoint
    itt = NT
    vmid = rt.record[itt - 1][1].cpu().numpy()
    A = 0.5 * (rt.ac[itt] + rt.ac[itt - 1])
    s_mid = dataclasses.replace(s, vlocal=vmid[None].copy())
    O = oracle_lib.Oracle(s_mid)
    O.update_kvector(A)
    ref, _, _ = O.taylor(A, snaps[-1].cpu().numpy())
    got = rt.psi_a.cpu().numpy()
    e = float(np.abs(got[..., : s.ny, :] - ref[..., : s.ny, :]).max() / np.abs(ref[..., : s.ny, :]).max())
    print(f"[parity] learned accepted step vs oracle Taylor: rel err {e:.3e}", flush=True)
    assert e <= 1e-12
