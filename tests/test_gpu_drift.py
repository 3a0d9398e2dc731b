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
