"""An RT run continued from a checkpoint written and read in SALMON's restart layout
(c1, 8 + 8 steps):

* against the oracle: the oracle runs the last 8 steps from the orbitals read back from the
  checkpoint files, with the field table of those steps -- the continued GPU run's current,
  dipole and electron count agree to the time-series bar 1e-9 (fixed potential and
  self-consistent predictor-corrector, single-file and sliced orbitals);
* against the uninterrupted run: bit for bit, also with trans_longi = 'lo', whose induced
  field continues from rtdata.bin (SALMON's write_rtdata / read_rtdata records)."""
import numpy as np
import pytest

from salmon2_amd import restart, system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _continued(s, tmp_path, sliced, **kw):
    from salmon2_amd.rt import RTPropagator

    full = RTPropagator(s, **kw)
    ref = full.run(16)
    full.close()
    first = RTPropagator(s, **kw)
    first.setup_outputs(16)
    for _ in range(8):
        first.step()
    first.checkpoint(str(tmp_path), sliced_block=sliced)
    first.close()
    cont = RTPropagator.from_restart(s, str(tmp_path), 16, **kw)
    assert cont.itt == 8
    for _ in range(8):
        cont.step()
    res = cont.results()
    cont.close()
    return ref, res


@pytest.mark.parametrize("scf", [False, True])
@pytest.mark.parametrize("sliced", [0, 4])
def test_continue_from_checkpoint(torch_dev, oracle_lib, tmp_path, scf, sliced):
    s = system.config("c1", nt=16)
    print(f"[gpu] restart continuation scf={scf} sliced={sliced}", flush=True)
    ref, res = _continued(s, tmp_path, sliced, scf=scf, pc=scf)
    for k in ("curr", "dip", "rne"):
        d = float(np.abs(res[k][8:] - ref[k][8:]).max())
        assert d == 0.0, k
    # the oracle from the orbitals in the files, over the field table of steps 9 .. 16
    d = restart.read_restart(str(tmp_path), (s.nx, s.ny, s.nz), no=s.no, nk=s.nk, ly=s.ly)
    ac = s.pulse.table(s.dt, 16)[8:]
    O = oracle_lib.Oracle(s)
    if scf:
        _, cur, dip, rne, _ = O.rt_run_scf(True, nt=8, psi=d["psi"], ac=ac)
    else:
        _, cur, dip, rne = O.rt_run(nt=8, psi=d["psi"], ac=ac)
    e = max(rel(res["curr"][8:], cur), rel(res["dip"][8:], dip), rel(res["rne"][8:], rne))
    print(f"[parity] restart scf={scf} sliced={sliced}: continued run vs oracle from the restart files {e:.3e}",
          flush=True)
    assert e <= 1e-9


def test_continue_longitudinal_field(torch_dev, tmp_path):
    s = system.config("c1", nt=16)
    print("[gpu] restart continuation trans_longi='lo'", flush=True)
    ref, res = _continued(s, tmp_path, 0, trans_longi="lo")
    for k in ("curr", "dip", "rne", "ac_tot"):
        assert np.array_equal(res[k][8:], ref[k][8:]), k
