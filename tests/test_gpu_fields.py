"""GPU RT driver with the 'lo' field feedback against the oracle, and the plasma-oscillation closed form."""
import math

import numpy as np
import pytest

from salmon2_amd import system
from tests.test_oracle_fields import _electron_gas

pytestmark = pytest.mark.gpu

SERIES_TOL = 1e-9


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


@pytest.mark.parametrize("scf", [False, True])
def test_lo_run_c1(torch_dev, oracle_lib, scf, tmp_path):
    s = system.config("c1")
    nt = 40
    print(f"[gpu] rt c1 trans_longi=lo scf={scf} steps={nt}", flush=True)
    from salmon2_amd.rt import RTPropagator

    rt = RTPropagator(s, scf=scf, pc=scf, trans_longi="lo")
    res = rt.run(nt)
    O = oracle_lib.Oracle(s)
    _, cur, dip, rne, ac, _ = O.rt_run_full(scf=scf, pc=scf, trans_longi="lo", nt=nt)
    errs = dict(current=rel(res["curr"], cur), dipole=rel(res["dip"], dip), ac_tot=rel(res["ac_tot"], ac[1 : nt + 1]))
    for k, v in errs.items():
        print(f"[parity] c1-lo{'-scf' if scf else ''} {k}: rel err {v:.3e}", flush=True)
        assert v <= SERIES_TOL, k
    # the induced field is a real signal
    assert np.abs(ac[1 : nt + 1] - s.pulse.table(s.dt, nt)[1 : nt + 1]).max() > 1e-9
    rt.write_rt_data(tmp_path / "Si_rt.data", res)
    rows = (tmp_path / "Si_rt.data").read_text().splitlines()[7:]
    assert len(rows) == nt and abs(float(rows[-1].split()[15]) - res["curr"][-1][2]) <= 5e-15 * abs(res["curr"][-1][2])


def test_lo_plasma_oscillation_gpu(torch_dev):
    s, omega = _electron_gas()
    from salmon2_amd.rt import RTPropagator

    print("[gpu] electron gas plasma oscillation", flush=True)
    rt = RTPropagator(s, trans_longi="lo")
    res = rt.run(s.nt)
    n = float(s.rocc.sum()) / omega
    theta = math.acos(1.0 - 2.0 * math.pi * n * s.dt ** 2)
    itt = np.arange(1, s.nt + 1)
    a = s.pulse.e_impulse
    ref = a * np.cos(theta * (itt - 0.5)) / math.cos(theta / 2)
    err = float(np.abs(res["ac_tot"][:, 2] - ref).max() / a)
    print(f"[parity] plasma oscillation A_tot: rel err {err:.3e}", flush=True)
    assert err < 1e-12
