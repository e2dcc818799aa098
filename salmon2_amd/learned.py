"""Learned Hartree+XC midpoint predictor for the predictor-corrector step.

The network (salmon2_amd.unet, BF16 MFMA kernels) maps the last four local potentials
V(t), V(t-dt), V(t-2dt), V(t-3dt) to a correction of the linear extrapolation
V(t) + (V(t) - V(t-dt))/2 of the midpoint potential V(t + dt/2).  It is trained on
trajectories of the plain predictor-corrector method on the GPU (the target is the
averaged potential (V(t) + V[rho'])/2 of src/rt/time_evolution_step.f90:402-405) of
seeded systems of the workload's shape, each with its own orbitals and pulse (amplitude
and polarisation drawn per trajectory).  In a run the step propagated with
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


def seeded_orbitals(s, seed, device="cuda"):
    """Orthonormal random orbitals for a training system, drawn on the device from `seed`
    (complex normal elements, Gram-Schmidt with the hvol inner product via a Cholesky QR),
    in the padded layout of ``System.psi``.  Training systems only: the benchmark workloads'
    orbitals come from ``system.random_orbitals`` / ``device_orbitals``."""
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    nx, ny, nz = s.nx, s.ny, s.nz
    N = nx * ny * nz
    out = torch.zeros((s.nk, s.no, s.nspin, nz, s.ly, nx), dtype=torch.complex128, device=device)
    for ik in range(s.nk):
        for isp in range(s.nspin):
            zt = torch.complex(torch.randn((s.no, N), generator=g, dtype=torch.float64, device=device),
                               torch.randn((s.no, N), generator=g, dtype=torch.float64, device=device))
            G = (zt.conj() @ zt.transpose(0, 1)).cpu().numpy() * s.hvol
            L = np.linalg.cholesky(G)
            rinv_t = torch.from_numpy(np.ascontiguousarray(np.linalg.inv(L.conj().T).T)).to(device)
            out[ik, :, isp, :, :ny, :] = (rinv_t @ zt).reshape(s.no, nz, ny, nx)
    return out


def trajectory(sysm, psi, pulse, nt, precision="fp64"):
    """Run the plain predictor-corrector from psi with `pulse`; return the samples (history of
    four potentials -> midpoint potential) as lists of device float64 tensors."""
    s = dataclasses.replace(sysm, pulse=pulse)
    prop = RTPropagator(s, psi=psi, scf=True, pc=True, precision=precision, record=True)
    v0 = prop.vcur.clone()
    prop.setup_outputs(nt)
    for _ in range(nt):
        prop.step()
    hist = [v0] + [r[0] for r in prop.record]          # V_0 .. V_nt
    mids = [r[1] for r in prop.record]                  # midpoint of step itt = 1..nt
    out = [([hist[itt - 1], hist[itt - 2], hist[itt - 3], hist[itt - 4]], mids[itt - 1]) for itt in range(4, nt + 1)]
    prop.close()
    return out


def generate(sysm, pulses, nt, precision="fp64"):
    """Trajectories of one system (its own orbitals) under several pulses."""
    samples = []
    for pulse in pulses:
        samples += trajectory(sysm, sysm.psi, pulse, nt, precision)
    return samples


def measure_scales(samples, vpsl):
    d = torch.stack([(h[0] - h[1]).flatten() for h, _ in samples])
    a = torch.stack([(h[0] - vpsl).flatten() for h, _ in samples])
    t = torch.stack([(m - h[0] - 0.5 * (h[0] - h[1])).flatten() for h, m in samples])
    return tuple(max(float(x.std()), 1e-30) for x in (d, a, t))


def to_tensors(pred, samples, vpsl):
    """(features, target) float32 device tensors of each sample."""
    out = []
    for h, m in samples:
        x = torch.zeros((pred.npts, 4), dtype=torch.float32, device=vpsl.device)
        t = torch.zeros(pred.npts, dtype=torch.float32, device=vpsl.device)
        pred.net.features(h[0], h[1], h[2], h[3], vpsl, pred.s_diff, pred.s_abs, x)
        pred.net.target(m, h[0], h[1], pred.s_out, t)
        out.append((x, t))
    return out


def train(pred, data, epochs=4, batch=8, lr=1e-3, decay=0.5, seed=0, log=None):
    """Adam on the MSE of the scaled target over data = [(features, target)], the step size
    multiplied by `decay` after every epoch; returns the per-epoch mean loss."""
    net = pred.net
    rng = np.random.default_rng(seed)
    losses = torch.zeros(len(data), dtype=torch.float32, device=data[0][0].device)  # one slot per sample
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(data))
        for b0 in range(0, len(order), batch):
            idx = order[b0 : b0 + batch]
            for i in idx:
                net.train_sample(data[i][0], data[i][1], losses[i : i + 1])
            net.adam_step(lr * decay ** ep, len(idx))
            if log and (b0 // batch) % 512 == 511:
                log(f"epoch {ep}: {b0 + len(idx)} of {len(order)} samples")
        hist.append(float(losses.mean().item()))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4e}")
    return hist


def build(config="c2", ntraj=200, nt=100, epochs=4, seed=20261300, tol=2e-5, nscale=8, log=print, **over):
    """Training set and training: ntraj seeded systems of the workload's shape (system seed,
    orbital seed and pulse all drawn from `seed` + i; none is a benchmark workload's seed),
    each propagated nt steps by the plain FP64 predictor-corrector with its own pulse
    (amplitude uniform in 0.001..0.02 au, polarisation uniform on the sphere); scales from
    the first `nscale` trajectories.  Returns (predictor, loss history, sample count)."""
    base_sys = _system.config(config, host_orbitals=False, **over)
    base = base_sys.pulse if base_sys.pulse is not None else _system.Pulse(A0=0.005, omega=0.5, tw=base_sys.nt * base_sys.dt)
    pulses = trajectory_pulses(ntraj, seed, base=base)
    pred, data, pending, vpsl = None, [], [], None
    for i in range(ntraj):
        s = _system.config(config, host_orbitals=False, seed=seed + 1 + i, **over)
        psi = seeded_orbitals(s, seed + 100003 + i)
        smp = trajectory(s, psi, pulses[i], nt)
        del psi
        vp = torch.from_numpy(np.ascontiguousarray(s.vlocal[0])).cuda()
        if pred is None:
            pending.append((smp, vp))
            if len(pending) == min(nscale, ntraj):
                allp = [x for p, _ in pending for x in p]
                scales = measure_scales(allp, pending[0][1])
                pred = Predictor(s.nx, s.ny, s.nz, seed=seed, stream=torch.cuda.current_stream(), tol=tol,
                                 scales=scales)
                for p, v in pending:
                    data += to_tensors(pred, p, v)
                pending = []
                if log:
                    log(f"scales {scales} from {min(nscale, ntraj)} trajectories")
        else:
            data += to_tensors(pred, smp, vp)
        del smp
        if log and (i + 1) % 20 == 0:
            log(f"{i + 1} trajectories, {len(data)} samples")
    torch.cuda.empty_cache()
    hist = train(pred, data, epochs=epochs, log=log)
    return pred, hist, len(data)


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
    ap.add_argument("--ntraj", type=int, default=200)
    ap.add_argument("--nt", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--tol", type=float, default=2e-5)
    ap.add_argument("--nscale", type=int, default=8, help="trajectories the feature scales are measured on")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    t0 = time.time()
    pred, hist, n = build(a.config, ntraj=a.ntraj, nt=a.nt, epochs=a.epochs, tol=a.tol, nscale=a.nscale,
                          log=lambda m: print(f"[train {time.time() - t0:7.1f}s] {m}", flush=True))
    out = a.out or weights_path(a.config)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    pred.save(out)
    print(f"[train] saved {out}: {a.ntraj} trajectories x {a.nt} steps, {n} samples; loss {hist[0]:.4e} -> "
          f"{hist[-1]:.4e}", flush=True)
