"""Mass drift over 1000 steps against the FP64 oracle trajectory.

Lenia does not conserve mass, so drift is |M_path(t) - M_FP64(t)| / M_FP64(t)
with both trajectories started from the same seeded state (the error
protocol of the workload definition).  The world is the three-ring rule
(R=26, mu=0.26, sigma=0.036, dt=0.1) at 1024^2 with 64 blobs of 48^2: the
4096^2 rule at a size the CPU FP64 oracle can follow for 1000 steps in about
a minute.  Gate: <= 1% over the whole run.
"""
import numpy as np
import pytest

from auto_amd import Sim, workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

STEPS = 1000


@pytest.fixture(scope="module")
def fp64_trajectory_mass(O):
    d = W.rings3_4096(side=1024)
    khats = O.kernel_hats64(d)
    s = O.init_state(d, 7)
    mass = [float(s.sum(dtype=np.float64))]
    for _ in range(STEPS):
        s = O.step64(d, s, "fft", khats)
        mass.append(float(s.sum(dtype=np.float64)))
    return np.array(mass)


@pytest.mark.parametrize("path", ["fft_fp32", "direct_bf16"])
def test_mass_drift_1000_steps(fp64_trajectory_mass, path):
    d = W.rings3_4096(side=1024, path=path)
    sim = Sim(d, 7)
    mass = [sim.stats().mass[0]]
    for _ in range(STEPS // 50):
        sim.step(50)
        mass.append(sim.stats().mass[0])
    mass = np.array(mass)
    ref = fp64_trajectory_mass[::50]
    drift = np.abs(mass - ref) / ref
    print(f"{path}: mass drift vs FP64 max {drift.max():.3e} final {drift[-1]:.3e} (mass {mass[-1]:.1f} vs {ref[-1]:.1f})")
    assert drift.max() <= 0.01


def _one_step_errors(O, d, sim, steps, every):
    """Every `every` steps: one step from the current state vs the FP64
    oracle step from the same input, per channel (max-abs, live cells)."""
    out = []
    t = 0
    while t < steps:
        s = sim.read_field("state")
        sim.step(1)
        got = sim.read_field("state")
        want = O.step64(d, s, "fft")
        out.append([(float(np.abs(got[c].astype(np.float64) - want[c]).max()), int(np.count_nonzero(want[c])))
                    for c in range(d.channels)])
        sim.step(every - 1)
        t += every
    return out


@pytest.mark.parametrize("path", ["fft_fp32", "direct_bf16"])
def test_one_step_error_along_1000_steps(O, path):
    """The three-ring rule at 1024^2: one-step max-abs vs FP64 sampled every
    100 steps over 1000 steps stays <= 1e-3 on both paths, while the world
    stays alive."""
    d = W.rings3_4096(side=1024, path=path)
    sim = Sim(d, 7)
    errs = _one_step_errors(O, d, sim, 1000, 100)
    worst = max(e[0][0] for e in errs)
    print(f"{path}: one-step max-abs vs FP64 every 100 steps: {[f'{e[0][0]:.1e}' for e in errs]}")
    assert worst <= 1e-3
    assert all(e[0][1] > 0 for e in errs)


@pytest.mark.parametrize("path", ["fft_fp32", "direct_bf16"])
def test_multichannel_error_per_channel_along_500_steps(O, path):
    """The 3-channel, 9-kernel rule at 1024^2: per channel, one-step max-abs
    vs FP64 sampled every 50 steps over 500 steps stays <= 1e-3."""
    d = W.multichannel_1024(seed=1, path=path)
    sim = Sim(d, 1)
    errs = _one_step_errors(O, d, sim, 500, 50)
    per = [max(e[c][0] for e in errs) for c in range(d.channels)]
    live = [[e[c][1] for c in range(d.channels)] for e in errs]
    print(f"{path}: worst per-channel one-step max-abs vs FP64 {per}; live cells per sample {live}")
    assert max(per) <= 1e-3
