"""Batched step over uneven systems: packed persistent stencil, per-system kernels on side
streams, device time index (salmon_hip_batch_*, salmon2_amd/batch.py).

Every point's arithmetic is the same as in the single-system path (only the tiling of the
work differs), so the batched run reproduces each system's own run to round-off (bound
1e-13 relative); the autotuned shapes keep the results; systems are checked against the
oracle (1e-9, the time-series bar).  Two sets: six hand-picked uneven systems (one of them
spin-polarised, one with s+p+d projectors; these two and the smallest also against the
oracle), and the 24 seeded systems of the batched workload (grids 16..40 per axis, 8..128
atoms, 16..256 orbitals, 1..27 k-points), every one against its own single-system run and the
two smallest against the oracle.  Stepping past the field tables is refused on the host.

Graph capture of the batched step is switched off in the library (a capture once ended in
a host segmentation fault whose cause was not found); its replay test is skipped.
"""
import numpy as np
import pytest

from salmon2_amd import system

pytestmark = pytest.mark.gpu

NT = 12


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    from salmon2_amd import hip

    hip.load()
    return torch


def _systems():
    return [
        system.config("tiny", nt=NT),
        system.make_si_system(n=(16, 20, 24), no=6, nkgrid=(1, 1, 1), seed=21, nt=NT),
        system.make_si_system(ncell=(2, 2, 2), n=32, no=8, nkgrid=(1, 1, 1), seed=22, nt=NT),
        system.make_si_system(ncell=(1, 1, 1), n=(40, 36, 32), no=3, seed=23, nt=NT),
        # spin-polarised member: two channels with their own V_local
        system.make_si_system(n=(16, 14, 12), no=3, nkgrid=(1, 1, 2), seed=47, nspin=2, nt=NT),
        # s+p+d projectors (9 per atom): the packed KB launches take the largest member's template
        system.make_si_system(n=12, no=4, nkgrid=(1, 1, 2), seed=53, lmax=2, nt=NT),
    ]


def _singles(systems, psis, nt):
    from salmon2_amd.rt import RTPropagator

    out = []
    for s, psi in zip(systems, psis):
        p = RTPropagator(s, psi=psi)
        out.append(p.run(nt))
        p.close()
    return out


def _check(name, got, ref, tol):
    for i, (g, r) in enumerate(zip(got, ref)):
        for k in ("curr", "dip", "rne"):
            e = rel(g[k], r[k])
            print(f"[parity] batch {name} system {i} {k}: rel err {e:.3e}", flush=True)
            assert e <= tol, (name, i, k, e)


def _vs_oracle(oracle_lib, name, s, psi_host, got, nt):
    O = oracle_lib.Oracle(s)
    _, cur, dip, rne = O.rt_run(nt=nt, psi=psi_host)
    e = max(rel(got["curr"], cur), rel(got["dip"], dip), rel(got["rne"], rne))
    print(f"[parity] batch {name} vs oracle: rel err {e:.3e}", flush=True)
    assert e <= 1e-9


def test_batch_eager_autotune(torch_dev, oracle_lib):
    from salmon2_amd.batch import BatchRunner

    systems = _systems()
    singles = _singles(systems, [None] * len(systems), NT)
    print("[gpu] batch of 6 uneven systems, eager", flush=True)
    br = BatchRunner(systems, NT)
    br.run(NT)
    eager = br.results(NT)
    _check("eager", eager, singles, 1e-13)
    print("[gpu] batch autotune", flush=True)
    choice, ms = br.batch.autotune(reps=2)
    print(f"[gpu] batch autotune choice={choice} {ms:.4f} ms", flush=True)
    br.reset()
    br.run(NT)
    _check("autotuned", br.results(NT), singles, 1e-13)
    # the field tables cover NT steps: one more is refused before anything is launched
    from salmon2_amd.hip import SalmonHipError

    with pytest.raises(SalmonHipError, match="past the field tables"):
        br.run(1)
    br.close()
    _vs_oracle(oracle_lib, "tiny", systems[0], systems[0].psi, eager[0], NT)
    _vs_oracle(oracle_lib, "spin2", systems[4], systems[4].psi, eager[4], NT)
    _vs_oracle(oracle_lib, "spd", systems[5], systems[5].psi, eager[5], NT)


def test_batch_24_seeded_systems(torch_dev, oracle_lib):
    torch = torch_dev
    from salmon2_amd.batch import BatchRunner

    nt = 6
    systems = system.batch_systems(nt=nt, host_orbitals=False)
    psis = [system.device_orbitals(s, "cuda") for s in systems]
    upd = sum(s.nx * s.ny * s.nz * s.norb for s in systems)
    print(f"[gpu] batch of the 24 seeded systems ({upd / 1e6:.1f} M grid-point.orbitals), {nt} steps", flush=True)
    singles = _singles(systems, psis, nt)
    br = BatchRunner(systems, nt, psi=psis)
    br.run(nt)
    got = br.results(nt)
    br.close()
    _check("24", got, singles, 1e-13)
    order = sorted(range(len(systems)), key=lambda i: systems[i].nx * systems[i].ny * systems[i].nz * systems[i].norb)
    for i in order[:2]:
        _vs_oracle(oracle_lib, f"24/system {i}", systems[i], psis[i].cpu().numpy(), got[i], nt)
    del psis
    torch.cuda.empty_cache()


def test_batch_rejects_steps_past_the_field_tables(torch_dev):
    """The device field tables end at nt: a step past them is refused on the host (no launch)."""
    from salmon2_amd.batch import BatchRunner
    from salmon2_amd.hip import SalmonHipError

    nt = 3
    br = BatchRunner([system.config("tiny", nt=nt), system.config("c1", nt=nt, no=8)], nt)
    br.run(nt)
    with pytest.raises(SalmonHipError, match="past the field tables"):
        br.run(1)
    with pytest.raises(SalmonHipError, match="outside"):
        br.batch.set_time(nt + 2)
    br.reset()
    br.run(nt)
    got = br.results(nt)
    br.close()
    assert all(np.all(np.isfinite(r["curr"])) for r in got)


def test_batch_capture_reports_switched_off(torch_dev):
    from salmon2_amd.batch import BatchRunner
    from salmon2_amd.hip import SalmonHipError

    br = BatchRunner([system.config("tiny", nt=2)], 2)
    with pytest.raises(SalmonHipError, match="switched off"):
        br.batch.capture()
    br.close()


@pytest.mark.skip(reason="graph capture of the batched step is switched off: a stream capture of the step once "
                         "ended in a host segmentation fault whose cause was not found")
def test_batch_graph_replay():
    pass
