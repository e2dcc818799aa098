"""Host logic of the learned predictor's training set (no GPU): the seeded pulses follow the
specification (amplitude uniform in 0.001..0.02 au, polarisation a unit vector, everything else the
workload's pulse), are reproducible from the seed, and the training systems' seeds never equal the
benchmark workloads' own."""
import numpy as np

from salmon2_amd import learned, system


def test_trajectory_pulses_follow_the_specification():
    base = system.config("c2", no=4, nt=2, host_orbitals=False).pulse
    ps = learned.trajectory_pulses(200, 20261300, base=base)
    assert len(ps) == 200
    amps = np.array([p.A0 for p in ps])
    assert amps.min() >= 0.001 and amps.max() <= 0.02
    # uniform: the sample spreads over the interval (two-sided, loose)
    assert amps.min() < 0.002 and amps.max() > 0.019 and abs(amps.mean() - 0.0105) < 0.0015
    dirs = np.array([p.epdir for p in ps])
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-12)
    # random polarisation: every axis takes both signs and the mean direction is near zero
    assert np.all(dirs.min(axis=0) < -0.5) and np.all(dirs.max(axis=0) > 0.5)
    assert np.linalg.norm(dirs.mean(axis=0)) < 0.2
    for p in ps:
        assert (p.omega, p.tw, p.shape if hasattr(p, "shape") else None) == \
            (base.omega, base.tw, base.shape if hasattr(base, "shape") else None)
    again = learned.trajectory_pulses(200, 20261300, base=base)
    assert [p.A0 for p in again] == list(amps) and [p.epdir for p in again] == [p.epdir for p in ps]


def test_training_seeds_are_held_out():
    # build() draws system i from seed + 1 + i and its orbitals from seed + 100003 + i
    seed, ntraj = 20261300, 200
    used = set(range(seed + 1, seed + 1 + ntraj)) | set(range(seed + 100003, seed + 100003 + ntraj))
    bench_seeds = {20261203, 20261204, 20261206, 7}  # c1, c2 / c2k1, c4, tiny
    assert not used & bench_seeds
