"""Many independent systems propagated together (the batched form).

Each system has its own Handle (grid, orbitals, projectors, fixed V_local) and its
own pulse; the batch runs the fixed-potential RT step for all of them with one
persistent stencil kernel per halo class and per Taylor term, one packed launch
per phase for all systems (phase table, projections, fused KB passes, scatter,
density, current, dipole), and A(t) tables and the step counter on the device.
Launches are eager: the library's graph capture of the batched step is switched
off (salmon_hip_batch_capture reports it).  torch provides the device buffers and
the stream.
"""
from __future__ import annotations

import numpy as np
import torch

from .hip import Batch, Handle


class BatchRunner:
    def __init__(self, systems, nt, device="cuda:0", psi=None):
        """psi: optional initial orbitals per system (device tensors or host arrays), for
        systems built with host_orbitals=False."""
        self.systems = list(systems)
        self.psi_init = list(psi) if psi is not None else [None] * len(self.systems)
        self.nt = nt
        self.dev = torch.device(device)
        torch.cuda.set_device(self.dev)
        # a stream of its own; torch work on the batch's buffers is issued on it too
        self.stream = torch.cuda.Stream(self.dev)
        with torch.cuda.stream(self.stream):
            self._setup(nt)

    def _setup(self, nt):
        self.handles = [Handle(s, stream=self.stream) for s in self.systems]
        self.batch = Batch(self.handles, stream=self.stream)
        self.buf = []
        zc = None
        for i, s in enumerate(self.systems):
            src = self.psi_init[i] if self.psi_init[i] is not None else s.psi
            if src is None:
                raise ValueError(f"system {i}: no orbitals (pass psi)")
            if isinstance(src, torch.Tensor):
                psi_a = src.to(self.dev, dtype=torch.complex128).contiguous().clone()
            else:
                psi_a = torch.from_numpy(np.ascontiguousarray(src)).to(self.dev)
            b = dict(psi0=psi_a.clone(), psi_a=psi_a, psi_b=torch.zeros_like(psi_a), work=torch.zeros_like(psi_a),
                     rho=torch.zeros((s.nspin, s.nz, s.ny, s.nx), dtype=torch.float64, device=self.dev),
                     ac=torch.from_numpy(np.ascontiguousarray(s.pulse.table(s.dt, nt) if s.pulse else
                                                              np.zeros((nt + 2, 3)))).to(self.dev),
                     curr=torch.zeros((nt, 3 * s.nspin), dtype=torch.float64, device=self.dev),
                     dip=torch.zeros((nt, 4), dtype=torch.float64, device=self.dev))
            self.buf.append(b)
            self.batch.set_system(i, b["psi_a"], b["psi_b"], b["work"], b["rho"], b["ac"], nt, b["curr"], b["dip"])
            z = s.zc()
            if zc is not None and not np.array_equal(z, zc):
                raise ValueError("the batched step needs one Taylor order and time step for all systems")
            zc = z
        self.zc = zc
        self.batch.prepare(zc)

    def reset(self):
        with torch.cuda.stream(self.stream):
            for b in self.buf:
                b["psi_a"].copy_(b["psi0"])
            self.batch.set_time(1)

    def run(self, nsteps):
        for _ in range(nsteps):
            self.batch.step()

    def close(self):
        torch.cuda.synchronize(self.dev)
        self.batch.close()
        for h in self.handles:
            h.close()

    def results(self, nsteps):
        torch.cuda.synchronize(self.dev)
        out = []
        for s, b in zip(self.systems, self.buf):
            c = b["curr"][:nsteps].cpu().numpy().reshape(nsteps, s.nspin, 3).sum(axis=1)
            d = b["dip"][:nsteps].cpu().numpy()
            psi = b["psi_a"] if nsteps % 2 == 0 else b["psi_b"]
            out.append(dict(curr=c, dip=d[:, :3], rne=d[:, 3], psi=psi))
        return out
