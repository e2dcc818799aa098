"""Learned Hartree+XC midpoint predictor for the predictor-corrector step.

The network (salmon2_amd.unet, BF16 MFMA kernels) maps the last four local potentials
V(t), V(t-dt), V(t-2dt), V(t-3dt) to a correction of the linear extrapolation
V(t) + (V(t) - V(t-dt))/2 of the midpoint potential V(t + dt/2).  It is trained on
trajectories of the plain predictor-corrector method on the GPU (the target is the
averaged potential (V(t) + V[rho'])/2 of src/rt/time_evolution_step.f90:402-405),
with seeded pulse amplitudes and polarisations.  In a run the step propagated with
the prediction is accepted only when its self-consistency residual
max|V_pred - (V(t) + V(t+dt))/2| is within `tol` (salmon2_amd.rt.RTPropagator).

Features and target are scaled by fixed factors measured on the training data
(std of the step differences, of V - Vpsl, of the target correction).
"""
from __future__ import annotations

import dataclasses
import json

import numpy as np
import torch

from . import system as _system
from .rt import RTPropagator
from .unet import UNet


class Predictor:
    def __init__(self, nx, ny, nz, seed=0, stream=None, tol=2e-5, scales=(1.0, 1.0, 1.0)):
        self.net = UNet(nx, ny, nz, seed=seed, stream=stream)
        self.npts = nx * ny * nz
        self.tol = float(tol)
        self.s_diff, self.s_abs, self.s_out = (float(v) for v in scales)
        dev = torch.device("cuda")
        self.x = torch.zeros((self.npts, 4), dtype=torch.float32, device=dev)
        self.nn = torch.zeros(self.npts, dtype=torch.float32, device=dev)

    def predict(self, hist, vpsl, vpred):
        """vpred = midpoint potential from hist = [V(t), V(t-dt), V(t-2dt), V(t-3dt)] (device float64)."""
        self.net.features(hist[0], hist[1], hist[2], hist[3], vpsl, self.s_diff, self.s_abs, self.x)
        self.net.forward(self.x, self.nn)
        self.net.predict(hist[0], hist[1], self.nn, self.s_out, vpred)

    def save(self, path):
        w = self.net.get_params()
        meta = dict(shape=list(self.net.shape), tol=self.tol, scales=[self.s_diff, self.s_abs, self.s_out])
        np.savez(path, params=w, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))

    @classmethod
    def load(cls, path, stream=None):
        d = np.load(path, allow_pickle=False)
        meta = json.loads(bytes(d["meta"]).decode())
        nz, ny, nx = meta["shape"]
        p = cls(nx, ny, nz, stream=stream, tol=meta["tol"], scales=meta["scales"])
        p.net.set_params(d["params"])
        return p


def trajectory_pulses(n, seed, amp=(0.001, 0.02), base=None):
    """Seeded pulses: amplitude uniform in `amp`, polarisation uniform on the sphere."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        out.append(dataclasses.replace(base, A0=float(rng.uniform(*amp)), epdir=tuple(float(x) for x in v)))
    return out


def generate(sysm, pulses, nt, precision="fp64"):
    """Run the plain predictor-corrector for each pulse; return the samples (history of four
    potentials -> midpoint potential) as lists of device float64 tensors."""
    samples = []
    for pulse in pulses:
        s = dataclasses.replace(sysm, pulse=pulse)
        prop = RTPropagator(s, scf=True, pc=True, precision=precision, record=True)
        v0 = prop.vcur.clone()
        prop.setup_outputs(nt)
        for _ in range(nt):
            prop.step()
        hist = [v0] + [r[0] for r in prop.record]          # V_0 .. V_nt
        mids = [r[1] for r in prop.record]                  # midpoint of step itt = 1..nt
        for itt in range(4, nt + 1):
            samples.append(([hist[itt - 1], hist[itt - 2], hist[itt - 3], hist[itt - 4]], mids[itt - 1]))
        prop.h.close()
    return samples


def measure_scales(samples, vpsl):
    d = torch.stack([(h[0] - h[1]).flatten() for h, _ in samples])
    a = torch.stack([(h[0] - vpsl).flatten() for h, _ in samples])
    t = torch.stack([(m - h[0] - 0.5 * (h[0] - h[1])).flatten() for h, m in samples])
    return tuple(max(float(x.std()), 1e-30) for x in (d, a, t))


def train(pred, samples, vpsl, epochs=4, batch=8, lr=1e-3, seed=0, log=None):
    """Adam on the MSE of the scaled target; returns the per-epoch mean loss."""
    net = pred.net
    xs, ts = [], []
    for h, m in samples:
        x = torch.zeros((pred.npts, 4), dtype=torch.float32, device=vpsl.device)
        t = torch.zeros(pred.npts, dtype=torch.float32, device=vpsl.device)
        net.features(h[0], h[1], h[2], h[3], vpsl, pred.s_diff, pred.s_abs, x)
        net.target(m, h[0], h[1], pred.s_out, t)
        xs.append(x)
        ts.append(t)
    rng = np.random.default_rng(seed)
    losses = torch.zeros(len(xs), dtype=torch.float32, device=vpsl.device)  # one slot per sample, read per epoch
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(xs))
        for b0 in range(0, len(order), batch):
            idx = order[b0 : b0 + batch]
            for i in idx:
                net.train_sample(xs[i], ts[i], losses[i : i + 1])
            net.adam_step(lr, len(idx))
        hist.append(float(losses.mean().item()))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
    return hist


def build(config="c2", ntraj=6, nt=200, epochs=4, seed=20261210, tol=2e-5, log=print, **over):
    """Generate trajectories, measure scales, train; returns (predictor, loss history)."""
    s = _system.config(config, **over)
    base = s.pulse if s.pulse is not None else _system.Pulse(A0=0.005, omega=0.5, tw=s.nt * s.dt)
    pulses = trajectory_pulses(ntraj, seed, base=base)
    samples = generate(s, pulses, nt)
    vpsl = torch.from_numpy(np.ascontiguousarray(s.vlocal[0])).cuda()
    scales = measure_scales(samples, vpsl)
    pred = Predictor(s.nx, s.ny, s.nz, seed=seed, stream=torch.cuda.current_stream(), tol=tol, scales=scales)
    if log:
        log(f"{len(samples)} samples from {ntraj} trajectories x {nt} steps, scales {scales}")
    hist = train(pred, samples, vpsl, epochs=epochs, log=log)
    return pred, hist


WEIGHTS = {"c2": "predictor_c2.npz"}


def weights_path(config):
    import os

    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", WEIGHTS.get(config, f"predictor_{config}.npz"))


if __name__ == "__main__":
    import argparse
    import os
    import time

    ap = argparse.ArgumentParser(description="train the midpoint-potential predictor on plain PC trajectories")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--ntraj", type=int, default=6)
    ap.add_argument("--nt", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--tol", type=float, default=2e-5)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    t0 = time.time()
    pred, hist = build(a.config, ntraj=a.ntraj, nt=a.nt, epochs=a.epochs, tol=a.tol,
                       log=lambda m: print(f"[train {time.time() - t0:7.1f}s] {m}", flush=True))
    out = a.out or weights_path(a.config)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    pred.save(out)
    print(f"[train] saved {out}; loss {hist[0]:.4e} -> {hist[-1]:.4e}", flush=True)
