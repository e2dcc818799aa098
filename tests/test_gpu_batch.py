"""Batched step over uneven systems: packed persistent stencil, device time index, hipGraph replay.

Every point's arithmetic is the same as in the single-system path (only the tiling of
the work differs), so the batched run reproduces each system's own run to round-off
(bound 1e-13 relative); the graph replay reproduces the eager batched run; the
autotuned shapes keep the results; the tiny system is checked against the oracle.
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
    ]


@pytest.fixture(scope="module")
def singles(torch_dev):
    from salmon2_amd.rt import RTPropagator

    out = []
    for s in _systems():
        out.append(RTPropagator(s).run(NT))
    return out


def _check(name, got, ref, tol):
    for i, (g, r) in enumerate(zip(got, ref)):
        for k in ("curr", "dip", "rne"):
            e = rel(g[k], r[k])
            print(f"[parity] batch {name} system {i} {k}: rel err {e:.3e}", flush=True)
            assert e <= tol, (name, i, k, e)


def test_batch_eager_graph_autotune(torch_dev, oracle_lib, singles):
    from salmon2_amd.batch import BatchRunner

    print("[gpu] batch of 4 uneven systems, eager", flush=True)
    br = BatchRunner(_systems(), NT)
    br.run(NT)
    eager = br.results(NT)
    _check("eager", eager, singles, 1e-13)
    print("[gpu] batch graph replay", flush=True)
    br.reset()
    br.run(NT, graph=True)
    graph = br.results(NT)
    _check("graph", graph, eager, 1e-15)
    print("[gpu] batch autotune", flush=True)
    choice, ms = br.batch.autotune(reps=2)
    print(f"[gpu] batch autotune choice={choice} {ms:.4f} ms", flush=True)
    br.reset()
    br.run(NT)
    _check("autotuned", br.results(NT), singles, 1e-13)
    # the tiny system against the oracle
    s = _systems()[0]
    O = oracle_lib.Oracle(s)
    _, cur, dip, rne = O.rt_run(nt=NT)
    e = max(rel(eager[0]["curr"], cur), rel(eager[0]["dip"], dip), rel(eager[0]["rne"], rne))
    print(f"[parity] batch tiny vs oracle: rel err {e:.3e}", flush=True)
    assert e <= 1e-9
