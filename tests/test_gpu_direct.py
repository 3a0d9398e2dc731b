"""GPU parity of the BF16 MFMA direct path against the FP64 oracle.

Tolerances: the state is rounded to bf16 (8-bit mantissa) on the way into
the MFMA and the taps are rounded to bf16 and renormalised, so the potential
carries ~1e-5 rms / ~1e-4 worst-case error; the gate for a step is the
north-star bar of 1e-3 max-abs vs FP64 from the same input state.
"""
import numpy as np
import pytest

from auto_amd import Sim, workloads as W
from auto_amd._abi import INIT_ZERO, EngineError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,radius,beta", [
    (64, 64, 9.0, [1.0]),
    (128, 64, 13.0, [1.0]),
    (256, 256, 26.0, [1.0, 2 / 3, 1 / 3]),
    (512, 128, 31.0, [0.5, 1.0]),
    (1024, 1024, 20.0, [0.3, 1.0, 0.7]),
])
def test_direct_conv_matches_fp64(O, w, h, radius, beta):
    d = W.desc(w, h, [dict(radius=radius, beta=beta, mu=0.2, sigma=0.03)], 0.1, path="direct_bf16",
               init=INIT_ZERO)
    sim = Sim(d, 1)
    s = np.random.default_rng(w + h).random((1, h, w), dtype=np.float32)
    sim.write_field("state", s)
    got = sim.read_field("conv0")
    want = O.conv_fft64(d, s, 0)
    err = np.abs(got - want).max()
    rms = np.sqrt(np.mean((got - want) ** 2))
    print(f"{w}x{h} R={radius} direct conv max-abs {err:.3e} rms {rms:.3e}")
    # A uniform-random field is the worst case for bf16 state rounding: the
    # potential's error is a weighted mean of per-cell rounding errors.
    assert err < 6e-4 and rms < 1.5e-4
    assert np.array_equal(sim.read_field("state"), s)


def test_direct_reference_known_answer(O):
    """automata.rs:300-362 bar (< 0.002 vs FP64 direct) on the direct path."""
    d = W.lenia_reference(128, 64, path="direct_bf16")
    seed = next(s for s in range(1, 64) if (O.init_state(d, s) > 0).any())
    sim = Sim(d, seed)
    s0 = sim.read_field("state")
    sim.step(1)
    err = np.abs(sim.read_field("state") - O.step64(d, s0, "direct")).max()
    print(f"direct KAT max-abs {err:.3e}")
    assert err < 0.002 and err <= 1e-3


def test_direct_rings3_4096_one_step(O):
    d = W.rings3_4096(path="direct_bf16")
    sim = Sim(d, 11)
    sim.step(3)
    s = sim.read_field("state")
    sim.step(1)
    got = sim.read_field("state")
    want = O.step64(d, s, "fft")
    diff = got.astype(np.float64) - want
    err = np.abs(diff).max()
    rel = np.linalg.norm(diff) / np.linalg.norm(want)
    print(f"direct 4096^2 rings3 max-abs {err:.3e} rel-L2 {rel:.3e}")
    assert err <= 1e-3


def test_direct_multichannel(O):
    d = W.multichannel_1024(seed=4, side=256, path="direct_bf16")
    sim = Sim(d, 9)
    khats = O.kernel_hats64(d)
    s = sim.read_field("state")
    worst = 0.0
    for _ in range(3):
        sim.step(1)
        got = sim.read_field("state")
        worst = max(worst, np.abs(got - O.step64(d, s, "fft", khats)).max())
        s = got
    print(f"direct multichannel worst per-step max-abs {worst:.3e}")
    assert worst <= 1e-3
    for k in (0, 4, 8):
        u = sim.read_field(f"conv{k}")
        assert np.abs(u - O.conv_fft64(d, s, k)).max() < 6e-4


def test_direct_single_ring_256_steps(O):
    d = W.single_ring_256(path="direct_bf16")
    sim = Sim(d, 1)
    khats = O.kernel_hats64(d)
    s = sim.read_field("state")
    worst = 0.0
    for _ in range(20):
        sim.step(1)
        got = sim.read_field("state")
        worst = max(worst, np.abs(got - O.step64(d, s, "fft", khats)).max())
        s = got
    print(f"direct 256^2 R=13 worst per-step max-abs {worst:.3e}")
    assert worst <= 1e-3


def test_direct_deterministic():
    d = W.rings3_4096(path="direct_bf16", side=512)
    a, b = Sim(d, 5), Sim(d, 5)
    a.step(7)
    b.step(7)
    assert np.array_equal(a.read_field("state"), b.read_field("state"))


def test_direct_rejects_large_radius():
    d = W.desc(256, 256, [dict(radius=40.0, mu=0.2, sigma=0.03)], 0.1, path="direct_bf16")
    with pytest.raises(EngineError) as e:
        Sim(d, 1)
    assert e.value.code == 6


@pytest.mark.parametrize("w,h,radius", [(128, 64, 9.0), (512, 256, 26.0)])
def test_direct_three_term_conv_is_fp32_accurate(O, w, h, radius):
    """K_hi s_hi + K_hi s_lo + K_lo s_hi on BF16 MFMA reaches ~FP32 accuracy."""
    d = W.desc(w, h, [dict(radius=radius, beta=[1.0, 0.5], mu=0.2, sigma=0.03)], 0.1, path="direct_bf16_x3",
               init=INIT_ZERO)
    sim = Sim(d, 1)
    s = np.random.default_rng(w).random((1, h, w), dtype=np.float32)
    sim.write_field("state", s)
    err = np.abs(sim.read_field("conv0") - O.conv_fft64(d, s, 0)).max()
    print(f"{w}x{h} R={radius} three-term conv max-abs {err:.3e}")
    assert err < 5e-6


def test_direct_auto_precision_meets_gate_on_high_gain_rule(O):
    """sigma=0.01, dt=0.2 (growth slope x24): the auto mode must pick three
    terms and stay within 1e-3 per step; one forced term does not."""
    kern = [dict(radius=9.0, mu=0.2, sigma=0.01)]
    errs = {}
    for path in ("direct_bf16", "direct_bf16_x1"):
        d = W.desc(256, 256, kern, 0.2, path=path, init=INIT_ZERO)
        sim = Sim(d, 1)
        s = np.random.default_rng(5).random((1, 256, 256), dtype=np.float32) * 0.4
        sim.write_field("state", s)
        sim.step(1)
        errs[path] = float(np.abs(sim.read_field("state") - O.step64(d, s, "fft")).max())
    print(errs)
    assert errs["direct_bf16"] <= 1e-4
    assert errs["direct_bf16"] < errs["direct_bf16_x1"]
