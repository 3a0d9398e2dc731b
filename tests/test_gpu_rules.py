"""SmoothLife and MNCA on the FFT path (GPU), against the oracle.  Height
4096 takes the streamed column pass with both kernels, other heights the
generic one; widths 128..4096 take the fused row pass.

SmoothLife's transition is smooth, so every step is compared with the FP64
oracle step and the reference-FP32 restatement from the same input.  MNCA's
bands are discontinuous: its potentials are compared with FP64, its update
must equal the oracle's band rule applied to the GPU's own potentials
exactly, and the full step may differ from FP64 only where an FP64 potential
lies within 1e-5 of a band edge.
"""
import numpy as np
import pytest

from auto_amd import Sim, workloads as W

pytestmark = pytest.mark.gpu

EDGES = (np.float32([0.210, 0.220, 0.350, 0.500, 0.750, 0.850, 0.120, 0.150]),
         np.float32([0.100, 0.280, 0.430, 0.550]))


@pytest.mark.parametrize("rule,seed", [("smoothlife", 3), ("mnca", 8)])
def test_seed_bit_exact(O, rule, seed):
    d = W.smoothlife(512, 256) if rule == "smoothlife" else W.mnca(512, 256)
    assert np.array_equal(Sim(d, seed).read_field("state"), O.init_state(d, seed))


@pytest.mark.parametrize("w,h", [(256, 256), (4096, 128), (64, 4096), (64, 8192)])
def test_smoothlife_steps(O, w, h):
    d = W.smoothlife(w, h)
    sim = Sim(d, 5)
    khats = O.kernel_hats64(d)
    spectra = O.kernel_spectra32(d)
    s = sim.read_field("state")
    worst64 = worst32 = 0.0
    for _ in range(8):
        sim.step(1)
        got = sim.read_field("state")
        worst64 = max(worst64, np.abs(got - O.step64(d, s, "fft", khats)).max())
        worst32 = max(worst32, np.abs(got - O.step_ref32(d, s, spectra)).max())
        s = got
    print(f"smoothlife {w}x{h}: worst per-step max-abs vs FP64 {worst64:.3e}, vs reference FP32 {worst32:.3e}")
    assert worst64 < 1e-4 and worst32 < 1e-4
    assert 0 < s.mean() < 1


@pytest.mark.parametrize("w,h,rings", [(256, 256, None), (4096, 128, None), (64, 4096, None), (64, 8192, None),
                                        (256, 256, (10, 15.5, 2, 6))])
def test_mnca_steps(O, w, h, rings):
    d = W.mnca(w, h) if rings is None else W.mnca(w, h, rings=rings)
    sim = Sim(d, 6)
    khats = O.kernel_hats64(d)
    flips = 0
    for _ in range(6):
        s = sim.read_field("state")
        u = np.stack([sim.read_field("potential0"), sim.read_field("potential1")])
        u64 = O.potentials64(d, s, khats)
        assert np.abs(u - u64).max() < 1e-5
        sim.step(1)
        got = sim.read_field("state")
        # The update is exactly the band rule on the GPU's potentials.
        assert np.array_equal(got, O.update64(d, u.astype(np.float64), s))
        near = np.zeros(u.shape[1:], bool)
        for k in range(2):
            for e in EDGES[k]:
                near |= np.abs(u64[k] - e) < 1e-5
        want = O.update64(d, u64, s)
        diff = (got != want)[0]
        assert not np.any(diff & ~near)
        flips += int(np.sum(got != s))
    print(f"mnca {w}x{h} rings {rings}: {flips} cell changes over 6 steps")
    assert flips > 0


@pytest.mark.parametrize("rule,w,h", [("smoothlife", 512, 512), ("mnca", 512, 512), ("smoothlife", 128, 4096)])
def test_rules_resume_bitwise(rule, w, h):
    """Resuming from a saved state (first step transforms the state itself)
    equals the uninterrupted run (spectra carried by the previous step).
    Height 4096 runs the streamed two-kernel column pass."""
    d = W.smoothlife(w, h) if rule == "smoothlife" else W.mnca(w, h)
    a, b = Sim(d, 2), Sim(d, 2)
    a.step(3)
    b.write_field("state", a.read_field("state"))
    a.step(5)
    b.step(5)
    assert np.array_equal(a.read_field("state"), b.read_field("state"))
