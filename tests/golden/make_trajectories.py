"""Writes the oracle's RT trajectories of the named workloads as golden fixtures.

Run in a container where the oracle builds (`python tests/golden/make_trajectories.py
[case ...]`); each case writes tests/golden/traj_<case>.npz with the per-step current
J(t) (nt, 3), dipole (nt, 3) and electron count (nt,), the field table and a fingerprint
of the seeded inputs (so a test can tell a changed generator from a wrong kernel).
The GPU tests regenerate the same seeded inputs, run the HIP path and compare with
these series (tests/test_gpu_trajectory.py).

Cases (all 'tr', middlepoint, 4th-order Taylor, dt = 0.08):
  c2k1_fixv    64-atom Si, 32^3, 128 orbitals, Gamma only, fixed V_local, 1000 steps
  c2k1_pc      the same with the self-consistent potential and the predictor-corrector
  c2_fixv      the full 64-atom workload (2x2x2 k), fixed V_local, 1000 steps
  c4o16_fixv   216-atom Si, 48^3, 4 k (2x2x1), 16 orbitals per k, fixed V_local, 500 steps
  c4_fixv10    the full 216-atom workload (432 orbitals x 4 k), fixed V_local, 10 steps
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = {
    "c2k1_fixv": dict(config="c2k1", over={}, scf=False, nt=1000),
    "c2k1_pc": dict(config="c2k1", over={}, scf=True, nt=1000),
    "c2_fixv": dict(config="c2", over={}, scf=False, nt=1000),
    "c4o16_fixv": dict(config="c4", over=dict(no=16), scf=False, nt=500),
    "c4_fixv10": dict(config="c4", over=dict(nt=10), scf=False, nt=10),
}


def fingerprint(s, psi=None) -> np.ndarray:
    """Seeded-input fingerprint: the orbitals' total norm, two weighted sums over a strided
    sample of the orbital elements (numpy array or torch tensor), V_local and projector
    sums.  Compared at 1e-12 relative: the orbital QR rounds differently with another BLAS
    thread count or on a GPU."""
    psi = s.psi if psi is None else psi
    flat = psi.reshape(-1)[::9973]
    w = np.cos(0.37 * np.arange(flat.shape[0], dtype=np.float64))
    if not isinstance(flat, np.ndarray):  # torch tensor
        norm = float((psi.real ** 2 + psi.imag ** 2).sum()) * s.hvol
        flat = flat.cpu().numpy()
    else:
        norm = float((np.abs(psi) ** 2).sum() * s.hvol)
    return np.array([
        norm,
        float((w * flat.real).sum()),
        float((w * flat.imag).sum()),
        float(s.vlocal.sum()),
        float(np.abs(s.uv).sum()) if s.natom else 0.0,
        float(np.abs(s.rinv_uvu).sum()) if s.natom else 0.0,
    ])


def system_of(case: str, host_orbitals=True):
    from salmon2_amd import system

    c = CASES[case]
    return system.config(c["config"], host_orbitals=host_orbitals, **c["over"])


def path_of(case: str) -> str:
    return os.path.join(HERE, f"traj_{case}.npz")


def make(case: str) -> str:
    from oracle import oracle

    c = CASES[case]
    s = system_of(case)
    O = oracle.Oracle(s)
    t0 = time.perf_counter()
    if c["scf"]:
        _, cur, dip, rne, _ = O.rt_run_scf(True, nt=c["nt"])
    else:
        _, cur, dip, rne = O.rt_run(nt=c["nt"])
    el = time.perf_counter() - t0
    out = path_of(case)
    np.savez_compressed(out, curr=cur, dip=dip, rne=rne, ac=s.pulse.table(s.dt, c["nt"]), fp=fingerprint(s),
                        nt=c["nt"], scf=int(c["scf"]), oracle_seconds=el,
                        threads=int(os.environ.get("OMP_NUM_THREADS", "0") or os.cpu_count()))
    print(f"{case}: {c['nt']} steps in {el:.1f} s -> {os.path.relpath(out, ROOT)}", flush=True)
    return out


def refresh_fingerprint(case: str) -> None:
    """Rewrite the input fingerprint of an existing fixture (no oracle run); the 216-atom
    workload's orbitals are drawn through the Cholesky-QR generator (equal to rounding)."""
    p = path_of(case)
    d = dict(np.load(p, allow_pickle=False))
    c = CASES[case]
    if c["config"] == "c4" and "no" not in c["over"]:
        from salmon2_amd import system

        s = system_of(case, host_orbitals=False)
        d["fp"] = fingerprint(s, system.device_orbitals(s, "cpu"))
    else:
        d["fp"] = fingerprint(system_of(case))
    np.savez_compressed(p, **d)
    print(f"{case}: fingerprint {d['fp']}", flush=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    if args and args[0] == "--refresh-fp":
        for case in args[1:]:
            refresh_fingerprint(case)
    else:
        for case in args or list(CASES):
            make(case)
