"""Real-time propagation driver on the GPU (host-side mirror of SALMON's RT loop).

Mirrors ``main_tddft`` (src/rt/main_tddft.f90:89-130) and the periodic,
fixed-potential (``yn_fix_func = 'y'``), middlepoint, ``trans_longi = 'tr'``
branch of ``time_evolution_step`` (src/rt/time_evolution_step.f90:116-127,
146-170, 243-264): per step ``itt``

  vec_Ac = (Ac_tot(itt) + Ac_tot(itt-1)) / 2 ; update_kvector_nonlocalpt
  taylor(spsi_in -> spsi_out, tpsi work)
  calc_density(spsi_out) ; calc_current with vec_Ac = Ac_tot(itt) ; subdip
  swap(spsi_in, spsi_out)

With ``scf=True`` (``yn_fix_func = 'n'``) the local potential is rebuilt
from the new density every step, V_local = Vpsl + Vh[rho] + Vxc[rho]
(time_evolution_step.f90:182-207), starting from the potential of the initial
orbitals; ``system.vlocal`` is then Vpsl.  With ``pc=True`` the step is the
predictor-corrector (time_evolution_step.f90:141-143, 354-423): a predictor
Taylor step with V(t), V' from its density, V = (V(t) + V')/2, then the
corrector Taylor step from the same input orbitals.  The library's taylor
leaves its input untouched, so the orbital copy of the reference
(psi_tmp, :370-372, 410-412) is not needed.

``learned=Predictor`` (salmon2_amd.learned) replaces the predictor Taylor
step by the network's midpoint potential V_pred: one Taylor step with it, then
the self-consistency residual max|V_pred - (V(t) + V[rho'])/2| is compared with
the predictor's tolerance ON THE DEVICE (salmon_unet_gate).  The flag it sets
gates the corrector branch (salmon_hip_set_skip): when the step is rejected,
the attempt serves as the predictor of the predictor-corrector -- V_mid =
(V(t) + V[rho'])/2, then the corrector Taylor step from the same orbitals --
and when it is accepted those launches do nothing and the attempt's orbitals,
density and potential stand.  No host synchronisation in the step; accepted
and attempted counts and the residuals are read back once (predictor_stats).

``precision="mixed"`` runs the Taylor chain H^n psi in complex FP32 with the
projections, the Taylor sum, density and current in FP64.

Fields follow calc_emfields (src/rt/em_field.f90:24-70): ``trans_longi="tr"``
keeps Ac_tot = Ac_ext; ``"lo"`` adds the induced field Ac_ind(t+dt) =
2 Ac_ind(t) - Ac_ind(t-dt) - 4 pi J(t) dt^2 (one 3-double read-back of J per
step).  ``write_rt_data`` writes SYSname_rt.data as write_rt_data_3d does
(src/io/write.f90:686-790, atomic units).

``propagator="aetrs"`` follows time_evolution_half_step_etrs
(src/rt/time_evolution_step.f90:119-123, 140-141, 425-475): Taylor coefficients of
dt/2; a half step with V(t) and A = Ac_tot(itt), the potential extrapolated to
3 V - 3 V(t-dt) + V(t-2dt) on the device (salmon_hip_vlocal_extrapolate; the
identity with a fixed potential, so skipped there), A = Ac_ext(itt+1) +
2 Ac_ind(itt) - Ac_ind(itt-1), the second half step.  As in the reference the
predictor-corrector is not used with it.

Orbitals stay resident in HBM for the whole run; per-step observables are
written to device arrays and read back once at the end.  torch is used for
device allocation, the streams and events, and the elementwise V_mid - V of the
corrector shortcut.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .hip import Handle


class RTPropagator:
    def __init__(self, sysm, device="cuda:0", psi=None, scf=False, pc=False, precision="fp64", trans_longi="tr",
                 learned=None, record=False, fuse_current=True, propagator="middlepoint"):
        self.s = sysm
        self.dev = torch.device(device)
        torch.cuda.set_device(self.dev)
        self.stream = torch.cuda.current_stream(self.dev)
        self.h = Handle(sysm, stream=self.stream)
        self.h.set_precision(precision)
        src = sysm.psi if psi is None else psi
        if src is None:
            raise ValueError("no orbitals: pass psi (system built with host_orbitals=False)")
        if isinstance(src, torch.Tensor):
            self.psi_a = src.to(self.dev, dtype=torch.complex128).contiguous().clone()
        else:
            self.psi_a = torch.from_numpy(np.ascontiguousarray(src)).to(self.dev)
        self.psi_b = torch.zeros_like(self.psi_a)
        self.work = torch.zeros_like(self.psi_a)
        self.rho = torch.zeros((sysm.nspin, sysm.nz, sysm.ny, sysm.nx), dtype=torch.float64, device=self.dev)
        if propagator not in ("middlepoint", "aetrs"):
            raise ValueError(f"propagator {propagator!r}: 'middlepoint' or 'aetrs'")
        self.aetrs = propagator == "aetrs"
        self.zc = sysm.zc(aetrs=self.aetrs)
        self.itt = 0
        if trans_longi not in ("tr", "lo"):
            raise ValueError(f"trans_longi {trans_longi!r}: 'tr' or 'lo'")
        self.trans_longi = trans_longi
        self.scf = bool(scf)
        # with a fixed potential the corrector average is the identity; aetrs takes precedence over the
        # predictor-corrector (time_evolution_step.f90:140-144)
        self.pc = bool(pc) and self.scf and not self.aetrs
        if self.scf:
            # Vpsl is the spin-independent part (system.vlocal[0]); V_local(ispin) = Vpsl + Vh[rho_1 + rho_2]
            # + Vxc(ispin)[rho_1, rho_2] with nspin = 2 (spin-polarised PZ)
            self.h.set_poisson(sysm.al)
            self.vpsl = torch.from_numpy(np.ascontiguousarray(sysm.vlocal[0])).to(self.dev)
            vshape = (sysm.nspin,) + tuple(self.vpsl.shape) if sysm.nspin == 2 else tuple(self.vpsl.shape)
            self.vcur = torch.zeros(vshape, dtype=torch.float64, device=self.dev)
            self.h.density(self.psi_a, self.rho)
            self.h.update_potential(self.rho, self.vpsl, vout=self.vcur)
            self.vmid = torch.zeros_like(self.vcur)  # the predictor-corrector midpoint potential of the last step
            self.dv = torch.zeros_like(self.vcur)  # V_mid - V of the predictor (corrector from the predictor's H psi)
            if self.aetrs:  # V(t-dt), V(t-2dt) of the extrapolation, both the initial V at the start
                self.vold1 = self.vcur.clone()
                self.vold2 = self.vcur.clone()
        # learned predictor (salmon2_amd.learned.Predictor): needs the self-consistent predictor-corrector step
        self.learned = learned
        if learned is not None:
            if not self.pc:
                raise ValueError("the learned predictor replaces the predictor of the predictor-corrector "
                                 "(scf=pc=True, propagator='middlepoint')")
            if sysm.nspin != 1:
                raise ValueError("the learned predictor takes one potential channel (nspin = 1)")
            self.hist = [torch.zeros_like(self.vpsl) for _ in range(4)]  # V(t), V(t-dt), V(t-2dt), V(t-3dt)
            self.hist[0].copy_(self.vcur)
            self.nhist = 1
            self.vpred = torch.zeros_like(self.vpsl)
            # the next step's inference is queued on the main stream right after V(t+dt), beside the
            # current's kernels on their own stream (at the 64-atom workload 0.12 ms per step faster
            # than inference on a third stream, whose convolutions then compete with both)
            self.pred_ready = False
            self.vnew = torch.zeros_like(self.vpsl)
            # V_H and V_xc of the attempt's density, kept for the rejected branch's midpoint potential
            self.vh_att = torch.zeros_like(self.vpsl)
            self.vxc_att = torch.zeros_like(self.vpsl)
            self.resid = torch.zeros(1, dtype=torch.float64, device=self.dev)
            self.flag = torch.zeros(1, dtype=torch.int32, device=self.dev)  # 1: this step's attempt accepted
            self.stats = torch.zeros(2, dtype=torch.int32, device=self.dev)  # accepted, attempted
            self.resid_log = torch.zeros(0, dtype=torch.float64, device=self.dev)
        self.record = [] if record else None
        # the current's projection of the new orbitals also forms the next step's first projection
        # (salmon_hip_current_next) when the next field is known in advance ('tr')
        self.fuse_current = bool(fuse_current) and sysm.natom > 0 and not self.aetrs
        self._vpred_rec = None
        # the current of a predictor-corrector step runs on a second stream while the density and
        # the potential update run on the first (independent reads of the new orbitals); joined at
        # the end of the step, before the next phase-table update
        self.cur_stream = torch.cuda.Stream(self.dev)
        self.ev_psi = torch.cuda.Event()
        self.ev_cur = torch.cuda.Event()
        self.cur_async = False

    def setup_outputs(self, nt):
        s = self.s
        self.nt = nt
        self.ac_ext = s.pulse.table(s.dt, nt) if s.pulse is not None else np.zeros((nt + 2, 3))
        self.ac_ind = np.zeros_like(self.ac_ext)
        self.ac = self.ac_ext.copy()  # Ac_tot
        self.curr_t = torch.zeros((nt, s.nspin, 3), dtype=torch.float64, device=self.dev)
        self.dip_t = torch.zeros((nt, 4), dtype=torch.float64, device=self.dev)
        if self.learned is not None:
            self.stats.zero_()
            self.resid_log = torch.zeros(nt, dtype=torch.float64, device=self.dev)

    def step(self):
        """One RT step itt -> itt+1 (time_evolution_step)."""
        self.itt += 1
        itt = self.itt
        A = self.ac[itt] if self.aetrs else 0.5 * (self.ac[itt] + self.ac[itt - 1])
        self.h.update_kvector(A)
        if self.learned is not None and self.nhist >= 4:
            self._learned_step()
        else:
            if self.aetrs:
                self.h.taylor(self.zc, self.psi_a, self.psi_b, self.work)
                if self.scf:
                    self.h.vlocal_extrapolate(self.vold1, self.vold2, vout=self.vcur)
                self.h.update_kvector(self.ac_ext[itt + 1] + 2.0 * self.ac_ind[itt] - self.ac_ind[itt - 1])
                self.psi_a, self.psi_b = self.psi_b, self.psi_a
                self.h.taylor(self.zc, self.psi_a, self.psi_b, self.work)
            elif self.pc:
                self.h.taylor(self.zc, self.psi_a, self.psi_b, self.work)
                self.h.density(self.psi_b, self.rho)
                self.h.update_potential(self.rho, self.vpsl, vmix=self.vcur, vout=self.vmid)
                self._corrector(self.vcur)
                self._current_async(itt)
            else:
                self.h.taylor(self.zc, self.psi_a, self.psi_b, self.work)
            self.h.density(self.psi_b, self.rho)
            if self.scf:
                self.h.update_potential(self.rho, self.vpsl, vout=self.vcur)
        if self.learned is not None:
            self.hist = [self.hist[3]] + self.hist[:3]
            self.hist[0].copy_(self.vcur)
            self.nhist = min(self.nhist + 1, 4)
            if self.nhist >= 4:
                P = self.learned
                P.net.set_stream(self.stream)
                P.predict(self.hist, self.vpsl, self.vpred)
                self.pred_ready = True
        if self.record is not None:
            # (V(t+dt), the step's midpoint potential, the prediction the step tried or None)
            self.record.append((self.vcur.clone(), self.vmid.clone(), self._vpred_rec))
            self._vpred_rec = None
        if not self.cur_async:
            self._current(itt)
        self.h.dipole(self.rho, self.dip_t[itt - 1])
        if self.cur_async:
            self.stream.wait_event(self.ev_cur)
            self.cur_async = False
        if self.trans_longi == "lo" and itt + 1 < len(self.ac):
            j = self.curr_t[itt - 1].sum(dim=0).cpu().numpy()  # synchronises this step
            self.ac_ind[itt + 1] = 2.0 * self.ac_ind[itt] - self.ac_ind[itt - 1] - 4.0 * math.pi * j * self.s.dt ** 2
            self.ac[itt + 1] = self.ac_ext[itt + 1] + self.ac_ind[itt + 1]
        self.psi_a, self.psi_b = self.psi_b, self.psi_a

    def _current_async(self, itt):
        """The current of the orbitals just completed (psi_b), on the second stream."""
        self.ev_psi.record(self.stream)
        self.cur_stream.wait_event(self.ev_psi)
        self.h.set_stream(self.cur_stream)
        try:
            self._current(itt)
        finally:
            self.h.set_stream(self.stream)
        self.ev_cur.record(self.cur_stream)
        self.cur_async = True

    def _current(self, itt):
        """J of the step's new orbitals (psi_b) with Ac_tot(itt); with a field fixed in advance the
        same gather also projects psi_b under the next step's midpoint field."""
        if self.fuse_current and self.trans_longi == "tr" and itt + 1 < len(self.ac):
            a_next = 0.5 * (self.ac[itt + 1] + self.ac[itt])
            self.h.current_next(self.ac[itt], self.psi_b, self.curr_t[itt - 1], a_next)
        else:
            self.h.current(self.ac[itt], self.psi_b, self.curr_t[itt - 1])

    def _corrector(self, v_pred):
        """The corrector Taylor step with V_mid (the handle's current potential) after a
        predictor with v_pred: its first H application is the predictor's plus (V_mid - v_pred) psi."""
        torch.sub(self.vmid, v_pred, out=self.dv)
        self.h.taylor_reuse(self.zc, self.psi_a, self.psi_b, self.work, self.dv)

    def _learned_step(self):
        """A step from the network's midpoint potential, gated on the device (see the module notes)."""
        P = self.learned
        if self.pred_ready:
            self.pred_ready = False
        else:
            P.net.set_stream(self.stream)
            P.predict(self.hist, self.vpsl, self.vpred)
        if self.record is not None:
            self._vpred_rec = self.vpred.clone()
        self.h.set_vlocal_device(self.vpred)
        self.h.taylor(self.zc, self.psi_a, self.psi_b, self.work)
        self.h.density(self.psi_b, self.rho)
        # V[rho'] in its parts (the same kernels' arithmetic as update_potential), so that the
        # rejected branch forms (V(t) + V[rho'])/2 without a second Hartree solve of rho'
        self.h.hartree(self.rho, self.vh_att)
        self.h.xc_pz(self.rho, self.vxc_att)
        self.h.update_vlocal(self.vpsl, self.vh_att, self.vxc_att, vout=self.vnew)
        P.net.residual(self.vpred, self.vcur, self.vnew, self.resid)
        P.net.gate(self.resid, getattr(P, "tol_run", P.tol), self.flag, self.stats, self.resid_log)
        # rejected: the attempt is the predictor, V_mid = (V(t) + V[rho'])/2, then the corrector,
        # the density and V(t+dt); accepted: these launches do nothing
        self.h.set_skip(self.flag)
        try:
            self.h.update_vlocal(self.vpsl, self.vh_att, self.vxc_att, vmix=self.vcur, vout=self.vmid)
            self._corrector(self.vpred)
            self.h.density(self.psi_b, self.rho)
            self.h.update_potential(self.rho, self.vpsl, vout=self.vcur)
        finally:
            self.h.set_skip(None)
        # accepted: V(t+dt) = V[rho'] (the handle's V_local already), the step's midpoint = V_pred
        self.h.copy_if(self.flag, self.vcur, self.vnew)
        self.h.copy_if(self.flag, self.vmid, self.vpred)
        self._current_async(self.itt)

    def predictor_stats(self):
        """(accepted, attempted, median residual) of the learned attempts since setup_outputs."""
        torch.cuda.synchronize(self.dev)
        acc, att = (int(v) for v in self.stats.cpu().numpy())
        n = min(att, self.resid_log.numel())
        med = float(np.median(self.resid_log[:n].cpu().numpy())) if n else None
        return acc, att, med

    @property
    def accepted(self):
        return self.predictor_stats()[0]

    @property
    def attempted(self):
        return self.predictor_stats()[1]

    def run(self, nt=None):
        nt = self.s.nt if nt is None else nt
        self.setup_outputs(nt)
        for _ in range(nt):
            self.step()
        return self.results()

    def close(self):
        """Release the device handle (buffers go with the object)."""
        torch.cuda.synchronize(self.dev)
        self.h.close()

    def results(self):
        torch.cuda.synchronize(self.dev)
        curr_s = self.curr_t.cpu().numpy()  # (nt, nspin, 3): the current of each spin channel
        curr = curr_s.sum(axis=1) if self.s.nspin == 2 else curr_s[:, 0]
        dip = self.dip_t.cpu().numpy()
        dt = self.s.dt
        n = len(curr)
        e_ext = -(self.ac_ext[1 : n + 1] - self.ac_ext[0:n]) / dt
        e_tot = -(self.ac[1 : n + 1] - self.ac[0:n]) / dt
        return dict(psi=self.psi_a, curr=curr, curr_s=curr_s, dip=dip[:, :3], rne=dip[:, 3], ac_ext=self.ac_ext[1 : n + 1],
                    ac_tot=self.ac[1 : n + 1], e_ext=e_ext, e_tot=e_tot)

    def checkpoint(self, directory, sliced_block=0):
        """The current orbitals, occupations and step count in SALMON's restart layout
        (checkpoint_rt, src/io/checkpoint_restart.f90:231-275; salmon2_amd.restart), with
        Vh_stock.bin (the Hartree potential of the current density, both records) for the
        self-consistent step and rtdata.bin (the fields at itt and itt+1) for trans_longi =
        'lo'.  The learned predictor's history is not part of it."""
        from . import restart

        torch.cuda.synchronize(self.dev)
        vh = None
        if self.scf:
            v = torch.zeros_like(self.vpsl)
            self.h.hartree(self.rho[0] + self.rho[1] if self.s.nspin == 2 else self.rho, v)
            vh = (v.cpu().numpy(),) * 2
        restart.write_checkpoint(directory, self.psi_a.cpu().numpy(), self.s.ny, self.s.rocc, self.itt,
                                 sliced_block=sliced_block, vh_stock=vh)
        if self.trans_longi == "lo":
            t = self.itt
            if t + 1 >= len(self.ac):
                raise ValueError(f"checkpoint at step {t}: the field table ends at row {len(self.ac) - 1}")
            restart.write_rtdata(directory, self.ac[t:t + 2], self.ac_ind[t:t + 2], self.ac_ind[0])

    @classmethod
    def from_restart(cls, sysm, directory, nt, **kw):
        """Continue from a restart directory (restart_rt, src/io/checkpoint_restart.f90:277-306):
        orbitals and occupations from its files, the step count from info.bin, and for
        trans_longi = 'lo' the fields of rtdata.bin (read_rtdata :1959-1990); nt = the total
        number of steps of the run (outputs are indexed from step 1)."""
        import dataclasses

        from . import restart

        d = restart.read_restart(directory, (sysm.nx, sysm.ny, sysm.nz), nspin=sysm.nspin, no=sysm.no, nk=sysm.nk,
                                 ly=sysm.ly)
        p = cls(dataclasses.replace(sysm, rocc=d["rocc"]), psi=d["psi"], **kw)
        p.setup_outputs(nt)
        t = d["iter"]
        p.itt = t
        if p.trans_longi == "lo":
            ac_tot, ac_ind, ac_ind0 = restart.read_rtdata(directory)
            p.ac[t:t + 2] = ac_tot
            p.ac_ind[t:t + 2] = ac_ind
            p.ac_ind[0] = ac_ind0
        return p

    def write_rt_data(self, path, res=None):
        """SYSname_rt.data of the steps run so far (write_rt_data_3d, atomic units; spin = 'polarized'
        columns for nspin = 2)."""
        res = self.results() if res is None else res
        write_rt_data(path, self.s.dt, res)


def _fe(x: float, d: int = 15, e: int = 3) -> str:
    """Fortran E(d+8).dEe edit of x: sign, 0.ddd..., E, signed e-digit exponent."""
    if x == 0.0 or not math.isfinite(x):
        mant, ex = 0.0, 0
    else:
        ex = math.floor(math.log10(abs(x))) + 1
        mant = abs(x) / 10.0 ** ex
        if round(mant, d) >= 1.0:
            mant, ex = mant / 10.0, ex + 1
    digits = f"{mant:.{d}f}"[1:]  # ".ddd"
    s = ("-" if x < 0 else "") + "0" + digits + ("E-" if ex < 0 else "E+") + f"{abs(ex):0{e}d}"
    return s.rjust(d + 5 + e)


def write_rt_data(path, dt, res):
    """Header and rows of SYSname_rt.data (src/io/write.f90:700-790): time, Ac_ext, E_ext,
    Ac_tot, E_tot, then Jm (columns 14-16, spin = 'unpolarized') or, when res["curr_s"] holds two
    spin channels, Jm_u and Jm_d (columns 14-19, spin = 'polarized'); format
    (F16.8, 99(1X, E23.15E3))."""
    cs = res.get("curr_s")
    polarized = cs is not None and np.asarray(cs).shape[1] == 2
    lines = ["# Real time calculation: ", "# Ac_ext: External vector potential field",
             "# E_ext: External electric field", "# Ac_tot: Total vector potential field",
             "# E_tot: Total electric field"]
    names = ["Time", "Ac_ext_x", "Ac_ext_y", "Ac_ext_z", "E_ext_x", "E_ext_y", "E_ext_z", "Ac_tot_x", "Ac_tot_y",
             "Ac_tot_z", "E_tot_x", "E_tot_y", "E_tot_z"]
    if polarized:
        lines += ["# Jm_u: Matter current density for spin-up electrons",
                  "# Jm_d: Matter current density for spin-down electrons"]
        names += ["Jm_u_x", "Jm_u_y", "Jm_u_z", "Jm_d_x", "Jm_d_y", "Jm_d_z"]
    else:
        lines.append("# Jm: Matter current density (electrons)")
        names += ["Jm_x", "Jm_y", "Jm_z"]
    lines.append("#" + "".join(f" {i + 1}:{nm}[a.u.]" for i, nm in enumerate(names)))
    for i in range(len(res["curr"])):
        jm = list(cs[i][0]) + list(cs[i][1]) if polarized else list(res["curr"][i])
        vals = list(res["ac_ext"][i]) + list(res["e_ext"][i]) + list(res["ac_tot"][i]) + list(res["e_tot"][i]) + jm
        lines.append(f"{(i + 1) * dt:16.8f}" + "".join(" " + _fe(float(v)) for v in vals))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
