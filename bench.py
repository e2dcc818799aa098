"""Benchmark: RT steps/s of the real-time Kohn-Sham propagation on one MI355X.

Workload (default): 64-atom Si supercell (2x2x2 diamond cells, a = 10.26 bohr),
32^3 grid, 128 complex128 orbitals x 8 k-points (2x2x2 Monkhorst-Pack), s+p
Gaussian KB projectors, 4th-order Taylor, dt = 0.08 au, sin^2 A(t) pulse of
amplitude 0.005 au.  Synthetic seeded inputs, orbitals HBM-resident.

Forms timed on the same initial state (each for warmup + steps RT steps):
  fixedv-fp64  V_local fixed (yn_fix_func = 'y'): phase table, Taylor step
               (4 H applications + sum), density, current, dipole
  pc-fp64      the reference's self-consistent step with the predictor-corrector
               (yn_fix_func = 'n', yn_predictor_corrector = 'y'): two Taylor
               steps, two Hartree + PZ-LDA potential updates per step; plain FP64,
               the baseline of the metric
  pc-mixed     the same step with the Taylor chain in complex FP32 (sums FP64)
  learned-mixed  pc-mixed with the predictor Taylor step replaced by the learned
               midpoint potential (salmon2_amd/learned.py), accepted per step by its
               self-consistency residual, else the plain predictor-corrector step
  learned-mixed-reuse  the same, but a rejected attempt serves as the predictor
               (one corrector step instead of the plain predictor-corrector)
  learned-mixed-backoff  learned-mixed-reuse that, after 8 rejected attempts in a row,
               attempts once every 32 steps until one is accepted (plain
               predictor-corrector steps in between)
  learned-mixed-open   every predicted step accepted (no gate): the speed the learned
               predictor offers and the error it costs on this workload
The batched form (salmon2_amd/batch.py) runs separately on 24 seeded uneven systems
(system.batch_systems): graph-replayed batched steps against the same systems
advanced one after another; reported under "batch".
Each form's current, dipole and electron count are compared with pc-fp64's
over the whole run.  `value` is the steps/s of the fastest learned+mixed form whose
current stays within 1e-6 (relative, max over the run) of pc-fp64's.

Prints ONE JSON line.  The roofline entry is for the dominant kernel, the
stencil of the first and middle Taylor terms (k_stencilz) in pc-fp64, timed live with
start/stop events recorded by its own launches over the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RT steps/s: learned+mixed vs plain FP64 GPU at \u22641e-6 rel current error"
HBM_PEAK_GBS = 8000.0

FORMS = {
    "fixedv-fp64": dict(scf=False, pc=False, precision="fp64"),
    "pc-fp64": dict(scf=True, pc=True, precision="fp64"),
    "pc-mixed": dict(scf=True, pc=True, precision="mixed"),
    "learned-mixed": dict(scf=True, pc=True, precision="mixed", learned=True),
    "learned-mixed-reuse": dict(scf=True, pc=True, precision="mixed", learned=True, fallback="reuse"),
    "learned-mixed-backoff": dict(scf=True, pc=True, precision="mixed", learned=True, fallback="reuse",
                                  backoff=(8, 32)),
    "learned-mixed-open": dict(scf=True, pc=True, precision="mixed", learned=True, tol=float("inf")),
}
BASE_FORM = "pc-fp64"
# `value` is the fastest learned+mixed form whose current stays within CUR_BAR of pc-fp64's
HEADLINE_FORMS = ("learned-mixed", "learned-mixed-reuse", "learned-mixed-backoff", "learned-mixed-open")
CUR_BAR = 1e-6


def cpu_baseline(sysm, steps: int):
    """The oracle (C/OpenMP restatement of SALMON's CPU loops) on a bounded sample of the
    baseline form (self-consistent potential, predictor-corrector)."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    O = oracle.Oracle(sysm)
    t0 = time.perf_counter()
    O.rt_run_scf(True, nt=steps)
    dt = time.perf_counter() - t0
    return {
        "value": steps / dt,
        "unit": "RT steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} predictor-corrector RT step(s) of the same workload, initial potential included "
                  f"(oracle/salmon_oracle.c, OpenMP, {threads} threads)",
    }


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (profiles/<workload>_traffic.json; counter values scaled by a
    1 GiB copy kernel), or None when no summary for this kernel exists."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f).get("per_dispatch_bytes", {})
    e = d.get(kernel)
    return None if e is None else e["read_bytes"] + e["write_bytes"]


def load_predictor(config, log):
    """Committed weights of the predictor for this workload, else a small one trained now."""
    import torch

    from salmon2_amd import learned

    path = learned.weights_path(config)
    if os.path.exists(path):
        return learned.Predictor.load(path, stream=torch.cuda.current_stream()), f"loaded {os.path.relpath(path, ROOT)}"
    pred, hist = learned.build(config, ntraj=2, nt=60, epochs=3, log=None)
    log(f"[bench] no committed predictor weights: trained 2 x 60-step trajectories, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    return pred, "trained in bench.py (2 trajectories x 60 steps), outside the timed region"


def run_form(s, form, steps, warmup, prof, predictor=None):
    import numpy as np
    import torch

    from salmon2_amd.rt import RTPropagator

    kw = dict(FORMS[form])
    tol = kw.pop("tol", None)
    if kw.pop("learned", False):
        kw["learned"] = predictor
        predictor.tol_run = predictor.tol if tol is None else tol
    prop = RTPropagator(s, **kw)
    prop.setup_outputs(warmup + steps)
    for _ in range(warmup):
        prop.step()
    torch.cuda.synchronize()
    if prof:
        prop.h.prof_start(max_launches=32 * steps + 32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        prop.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern = {}
    if prof:
        for k in ("hpsi_taylor1", "hpsi_taylor", "hpsi_taylor_last", "proj", "proj_current", "nl_scatter"):
            ms, n = prop.h.prof_read(k)
            kern[k] = {"avg_ms": ms / max(n, 1), "launches": n}
        prop.h.prof_end()
    res = prop.results()
    res["psi"] = None
    if prop.learned is not None:
        res["accepted"] = prop.accepted
        res["attempted"] = prop.attempted
        res["resid_median"] = float(np.median(prop.resid_log)) if prop.resid_log else None
    assert np.all(np.isfinite(res["curr"])), f"non-finite current in {form}"
    prop.h.close()
    del prop
    torch.cuda.empty_cache()
    return dict(elapsed=t1 - t0, kern=kern, **res)


def run_batch(steps, warmup, log):
    """Graph-replayed batched steps of the seeded uneven systems, and the same systems one by one."""
    import numpy as np
    import torch

    from salmon2_amd import system
    from salmon2_amd.batch import BatchRunner
    from salmon2_amd.rt import RTPropagator

    systems = system.batch_systems(nt=warmup + steps)
    upd = sum(s.nx * s.ny * s.nz * s.norb for s in systems)
    br = BatchRunner(systems, warmup + steps)
    choice, tune_ms = br.batch.autotune(reps=2)
    br.reset()
    br.run(warmup, graph=False)
    br.batch.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    br.batch.replay(steps // 2)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nb = 2 * (steps // 2)
    res = br.results(warmup + nb)
    bad = [i for i, r in enumerate(res) if not np.all(np.isfinite(r["curr"]))]
    assert not bad, f"non-finite current in batched systems {bad}"
    del br
    torch.cuda.empty_cache()
    # the same systems advanced one after another (single-system path, eager)
    seq_steps = max(2, steps // 10)
    tseq = 0.0
    for s in systems:
        p = RTPropagator(s)
        p.setup_outputs(seq_steps + 3)
        for _ in range(3):
            p.step()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(seq_steps):
            p.step()
        torch.cuda.synchronize()
        tseq += time.perf_counter() - a
        p.h.close()
        del p
    torch.cuda.empty_cache()
    log(f"[bench] batch: {nb / (t1 - t0):.2f} batched steps/s, one-by-one {seq_steps / tseq:.2f}")
    return {
        "systems": len(systems),
        "grid_point_orbital_updates_per_step": upd,
        "steps": nb,
        "batched_steps_per_s": nb / (t1 - t0),
        "grid_point_orbital_updates_per_s": upd * nb / (t1 - t0),
        "one_by_one_steps_per_s": seq_steps / tseq,
        "speedup_vs_one_by_one": (nb / (t1 - t0)) / (seq_steps / tseq),
        "autotune_choice": choice,
        "form": "fixed-potential FP64 step; one persistent stencil kernel per halo class and Taylor term, "
                "two steps per hipGraph replay",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--forms", default=",".join(FORMS))
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--batch-steps", type=int, default=200, help="batched-form steps (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the run, copy 1 GiB device-to-device (HBM-traffic counter calibration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU path is not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from salmon2_amd import system

    s = system.config(args.config)
    forms = [f for f in args.forms.split(",") if f]
    if BASE_FORM not in forms:
        forms.insert(0, BASE_FORM)
    runs = {}
    predictor, pred_note = (None, None)
    if any(FORMS[f].get("learned") for f in forms):
        predictor, pred_note = load_predictor(args.config, lambda m: print(m, file=sys.stderr, flush=True))
    for f in forms:
        runs[f] = run_form(s, f, args.steps, args.warmup, prof=(f == BASE_FORM), predictor=predictor)
        print(f"[bench] {f}: {args.steps / runs[f]['elapsed']:.2f} steps/s", file=sys.stderr, flush=True)

    npts = s.nx * s.ny * s.nz
    norb = s.norb
    base = runs[BASE_FORM]

    def rel(a, b):
        return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())

    form_out = {}
    for f, r in runs.items():
        sps = args.steps / r["elapsed"]
        form_out[f] = {
            "steps_per_s": sps,
            "ms_per_step": 1e3 * r["elapsed"] / args.steps,
            "speedup_vs_" + BASE_FORM: sps / (args.steps / base["elapsed"]),
            "grid_point_orbital_updates_per_s": sps * npts * norb,
        }
        if "accepted" in r:
            form_out[f]["predictor"] = {"accepted_steps": r["accepted"], "attempted_steps": r["attempted"],
                                        "median_residual": r["resid_median"],
                                        "tol": FORMS[f].get("tol", predictor.tol), "weights": pred_note}
        if f != BASE_FORM and FORMS[f]["scf"]:
            form_out[f]["rel_err_vs_" + BASE_FORM] = {
                "current": rel(r["curr"], base["curr"]),
                "dipole": rel(r["dip"], base["dip"]),
                "electron_count": float(np.abs(r["rne"] - base["rne"]).max() / base["rne"][0]),
            }
    ok = [f for f in HEADLINE_FORMS if f in form_out and form_out[f]["rel_err_vs_" + BASE_FORM]["current"] <= CUR_BAR]
    head = max(ok, key=lambda f: form_out[f]["steps_per_s"]) if ok else BASE_FORM
    kern = base["kern"]
    # algorithmic bytes of one launch of the terms without the Taylor sum (first and middle terms, the
    # same kernel and work): read H^(n-1) psi (16 B) and write H^n psi (16 B) per point-orbital, plus
    # V_local once (8 B per point); the average duration is over all those launches in the timed region
    alg_bytes = 32.0 * npts * norb + 8.0 * npts
    k1, km = kern["hpsi_taylor1"], kern["hpsi_taylor"]
    nl = k1["launches"] + km["launches"]
    avg_ms = (k1["avg_ms"] * k1["launches"] + km["avg_ms"] * km["launches"]) / max(nl, 1)
    achieved = alg_bytes / (avg_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": form_out[head]["steps_per_s"],
        "unit": "RT steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": form_out[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if FORMS[head]["precision"] == "fp64" else "c64 chain / f64 sums",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: Si {s.natom} atoms, {s.nx}x{s.ny}x{s.nz} grid, {s.no} orbitals x {s.nk} k, "
                        f"{s.nlma} KB projectors, Taylor order {s.n_hamil}, dt {s.dt}",
            "form": head,
            "form_rule": f"fastest of {list(HEADLINE_FORMS)} with current within {CUR_BAR:g} of {BASE_FORM}",
            "baseline_form": BASE_FORM,
            "grid_point_orbital_updates_per_s": form_out[head]["grid_point_orbital_updates_per_s"],
        },
        "forms": form_out,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_stencilz (first and middle Taylor terms of pc-fp64: 25-point stencil + V + k.grad on y "
                      "tiles with merged halo rows and the z part formed 4 planes late, LDS-DMA fed)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.config, "k_stencilz<d2,double,d2>"),
            "alg_bytes_per_launch": alg_bytes,
            "avg_launch_ms": avg_ms,
            "launches": nl,
        },
        "kernels_ms": {k: round(v["avg_ms"], 5) for k, v in kern.items()},
    }
    runs = None
    torch.cuda.empty_cache()
    if args.batch_steps > 0:
        out["batch"] = run_batch(args.batch_steps, 4, lambda m: print(m, file=sys.stderr, flush=True))
    if args.calib_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = src.numel() * 8
        del src, dst
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
