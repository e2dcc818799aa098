"""An RT run continued from a checkpoint written and read in SALMON's restart layout
reproduces the uninterrupted run: the orbitals round-trip bit for bit through the files,
so the continued current, dipole and electron count equal the uninterrupted ones exactly
(fixed potential and self-consistent predictor-corrector, c1, 8 + 8 steps)."""
import numpy as np
import pytest

from salmon2_amd import system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


@pytest.mark.parametrize("scf", [False, True])
@pytest.mark.parametrize("sliced", [0, 4])
def test_continue_from_checkpoint(torch_dev, tmp_path, scf, sliced):
    from salmon2_amd.rt import RTPropagator

    s = system.config("c1", nt=16)
    print(f"[gpu] restart continuation scf={scf} sliced={sliced}", flush=True)
    full = RTPropagator(s, scf=scf, pc=scf)
    ref = full.run(16)
    first = RTPropagator(s, scf=scf, pc=scf)
    first.setup_outputs(16)
    for _ in range(8):
        first.step()
    first.checkpoint(str(tmp_path), sliced_block=sliced)
    cont = RTPropagator.from_restart(s, str(tmp_path), 16, scf=scf, pc=scf)
    assert cont.itt == 8
    for _ in range(8):
        cont.step()
    res = cont.results()
    for k in ("curr", "dip", "rne"):
        d = float(np.abs(res[k][8:] - ref[k][8:]).max())
        print(f"[parity] restart scf={scf} sliced={sliced} {k}: max abs diff {d:.3e}", flush=True)
        assert d == 0.0, k
