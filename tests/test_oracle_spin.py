"""Pins the oracle's spin-polarised (collinear, nspin = 2) path against its spin-unpolarised one.

The orbitals of the two spin channels are independent blocks of the same Hamiltonian up to the
channel's V_local (hamiltonian.f90: V_local(ispin) in the local term; the kinetic, k + A and
nonlocal terms do not depend on the spin).  With both channels holding the nspin = 1 orbitals:

* hpsi / one Taylor step of channel s equal the nspin = 1 ones with V_local = V_s (a wrong
  channel-to-potential index fails this: the two potentials differ);
* the density of each channel with occupation 1 equals half the nspin = 1 density with
  occupation 2 (exactly: the factor 2 is a power of two), the dipole of the two channels
  together equals the nspin = 1 dipole;
* the current of each channel is half the nspin = 1 current (the current does not see V).
"""
import dataclasses

import numpy as np

from salmon2_amd import system


def _pair(**kw):
    s2 = system.make_si_system(nspin=2, **kw)
    s1 = system.make_si_system(nspin=1, **kw)
    psi2 = np.repeat(s1.psi, 2, axis=2)  # (nk, no, nspin, nz, ly, nx): both channels = the nspin = 1 orbitals
    s2 = dataclasses.replace(s2, psi=psi2)
    return s1, s2


def test_spin_channels_follow_their_potentials(oracle_lib):
    s1, s2 = _pair(n=(12, 10, 12), no=3, nkgrid=(1, 1, 2), seed=41)
    assert s2.nspin == 2 and not np.array_equal(s2.vlocal[0], s2.vlocal[1])
    assert np.array_equal(s2.vlocal[0], s1.vlocal[0])
    A = np.array([0.012, -0.008, 0.02])
    O2 = oracle_lib.Oracle(s2)
    O2.update_kvector(A)
    h2 = O2.hpsi(A, s2.psi)
    t2, _, _ = O2.taylor(A, s2.psi)
    for sp in range(2):
        s1v = dataclasses.replace(s1, vlocal=s2.vlocal[sp][None].copy())
        O1 = oracle_lib.Oracle(s1v)
        O1.update_kvector(A)
        h1 = O1.hpsi(A, s1.psi)
        t1, _, _ = O1.taylor(A, s1.psi)
        assert np.array_equal(h2[:, :, sp:sp + 1], h1), sp
        assert np.array_equal(t2[:, :, sp:sp + 1], t1), sp
    # the channels really differ where V does
    assert np.abs(h2[:, :, 0] - h2[:, :, 1]).max() > 1e-3


def test_spin_density_current_dipole(oracle_lib):
    s1, s2 = _pair(n=12, no=4, nkgrid=(1, 1, 2), seed=43)
    A = np.array([0.02, 0.0, -0.01])
    O1 = oracle_lib.Oracle(s1)
    O2 = oracle_lib.Oracle(s2)
    O1.update_kvector(A)
    O2.update_kvector(A)
    rho1 = O1.density(s1.psi)
    rho2 = O2.density(s2.psi)
    assert rho2.shape == (2, s1.nz, s1.ny, s1.nx)
    assert np.array_equal(rho2[0], 0.5 * rho1[0]) and np.array_equal(rho2[1], 0.5 * rho1[0])
    j1 = O1.current(A, s1.psi)
    j2 = O2.current(A, s2.psi)
    assert np.abs(j2[0] - 0.5 * j1[0]).max() <= 1e-15 * np.abs(j1).max()
    assert np.abs(j2[1] - 0.5 * j1[0]).max() <= 1e-15 * np.abs(j1).max()
    d1 = O1.dipole(rho1)
    d2 = O2.dipole(rho2)
    assert np.abs(d2 - d1).max() <= 1e-14 * np.abs(d1).max()
