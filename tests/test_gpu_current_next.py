"""salmon_hip_current_next: the current's gather of the new orbitals also forms the next step's
phase table and first projection.  The RT runs with it must equal the runs without it bit for
bit (same kernels' arithmetic, same summation order), in every form that uses it; a field other
than the one built ahead discards the prepared table and projection."""
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


def _run(s, nt, fuse, **kw):
    from salmon2_amd.rt import RTPropagator

    p = RTPropagator(s, fuse_current=fuse, **kw)
    r = p.run(nt)
    psi = p.psi_a.cpu().numpy()
    p.close()
    return r, psi


@pytest.mark.parametrize("name,kw", [
    ("tiny", dict()),
    ("c1", dict()),
    ("c1", dict(scf=True, pc=True)),
    ("c1", dict(scf=True, pc=True, precision="mixed")),
    ("c2", dict(no=16, nt=12)),
])
def test_fused_current_equals_unfused(torch_dev, name, kw):
    kw = dict(kw)
    over = {k: kw.pop(k) for k in ("no", "nt") if k in kw}
    s = system.config(name, **over)
    nt = min(s.nt, 12)
    print(f"[gpu] current_next {name} {kw} {nt} steps", flush=True)
    r0, p0 = _run(s, nt, False, **kw)
    r1, p1 = _run(s, nt, True, **kw)
    d = max(float(np.abs(r0[k] - r1[k]).max()) for k in ("curr", "dip", "rne"))
    print(f"[parity] current_next {name} {kw}: max |diff| of J, dipole, N_e {d:.3e}; orbitals "
          f"{float(np.abs(p0 - p1).max()):.3e}", flush=True)
    for k in ("curr", "dip", "rne"):
        assert np.array_equal(r0[k], r1[k]), k
    assert np.array_equal(p0, p1)


def test_other_field_discards_prepared_projection(torch_dev):
    torch = torch_dev
    from salmon2_amd.hip import Handle

    s = system.config("tiny")
    h = Handle(s, stream=torch.cuda.current_stream())
    psi = torch.from_numpy(s.psi).cuda()
    cur = torch.zeros(3 * s.nspin, dtype=torch.float64, device="cuda")
    A0, A1, A2 = np.array([0.01, 0.0, 0.0]), np.array([0.02, -0.01, 0.0]), np.array([-0.03, 0.01, 0.02])
    outs = []
    for prepared in (False, True):
        h.update_kvector(A0)
        if prepared:
            h.current_next(A0, psi, cur, A1)  # prepares A1's table and psi's projection under it
        h.update_kvector(A2)                   # another field: both discarded
        out, work = torch.zeros_like(psi), torch.zeros_like(psi)
        h.taylor(s.zc(), psi, out, work)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    h.close()
    assert np.array_equal(outs[0], outs[1])
