"""GPU parity of the BF16 MFMA direct path against the FP64 oracle.

Tolerances: the state is rounded to bf16 (8-bit mantissa) on the way into
the MFMA and the taps are rounded to bf16 and renormalised, so the potential
carries ~1e-5 rms / ~1e-4 worst-case error; the gate for a step is the
1e-3 max-abs bar vs FP64 from the same input state.
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
    got = sim.read_field("potential0")
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
        u = sim.read_field(f"potential{k}")
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
    err = np.abs(sim.read_field("potential0") - O.conv_fft64(d, s, 0)).max()
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


VARIANTS = [(128, 128, 0), (128, 64, 0), (64, 64, 0), (128, 64, 1), (64, 64, 1), (128, 128, 2), (128, 64, 2),
            (64, 64, 2)]


def _run_variants(d, s, steps):
    """State after `steps` steps from s under every admissible variant."""
    out = {}
    for v in VARIANTS:
        sim = Sim(d, 1)
        try:
            sim.set_direct_tile(*v)
        except EngineError:
            continue
        assert sim.direct_tile() == v
        sim.write_field("state", s)
        sim.step(steps)
        out[v] = sim.read_field("state")
    return out


@pytest.mark.parametrize("w,h,radius", [(64, 64, 9.0), (256, 128, 13.0), (2048, 2048, 26.0), (1024, 4096, 31.0)])
def test_direct_tile_variants_agree_bitwise(O, w, h, radius):
    """Every admissible kernel variant (three tile shapes and the persistent
    ring kernel at two widths) runs the same MFMA sequence per cell, so steps
    from the same state give the same bits.  At 2048^2 and 1024x4096 a ring
    workgroup walks 2-4 tiles, so the halo rows it carries from one tile to
    the next (and the rows it streams in during the MFMAs) must equal what the
    tile kernel restages.  The ring kernel's step is also checked against FP64."""
    d = W.desc(w, h, [dict(radius=radius, beta=[1.0], mu=0.15, sigma=0.015)], 0.1, path="direct_bf16_x1",
               init=INIT_ZERO)
    s = np.random.default_rng(w + 3 * h).random((1, h, w), dtype=np.float32)
    got = _run_variants(d, s, 4)
    assert (128, 64, 0) in got or (64, 64, 0) in got
    assert not [v for v in got if v[2] == 2], "one kernel has no split variant"
    rings = [v for v in got if v[2] == 1]
    assert rings, "the ring kernel is admissible for one channel, one kernel, one term"
    ref = next(iter(got.values()))
    for v, g in got.items():
        assert np.array_equal(g, ref), v
    sim = Sim(d, 1)
    sim.set_direct_tile(*rings[0])
    sim.write_field("state", s)
    sim.step(1)
    err = np.abs(sim.read_field("state") - O.step64(d, s, "fft")).max()
    print(f"{w}x{h} R={radius}: variants {sorted(got)}; ring step max-abs vs FP64 {err:.3e}")
    assert err <= 1e-3


def test_direct_ring_4096_matches_tile_kernel():
    """4096^2 three-ring world: ring workgroups walk 8 tiles each."""
    d = W.rings3_4096(path="direct_bf16")
    a, b = Sim(d, 5), Sim(d, 5)
    a.set_direct_tile(128, 64, 0)
    b.set_direct_tile(128, 64, 1)
    a.step(3)
    b.step(3)
    assert np.array_equal(a.read_field("state"), b.read_field("state"))


def test_direct_ring_not_admissible_for_three_terms_or_channels():
    d = W.desc(256, 256, [dict(radius=13.0, beta=[1.0], mu=0.15, sigma=0.015)], 0.1, path="direct_bf16_x3",
               init=INIT_ZERO)
    sim = Sim(d, 1)
    with pytest.raises(EngineError):
        sim.set_direct_tile(128, 64, 1)
    assert sim.direct_tile()[2] == 0
    m = Sim(W.multichannel_1024(seed=4, side=256, path="direct_bf16"), 2)
    with pytest.raises(EngineError):
        m.set_direct_tile(128, 64, 1)


@pytest.mark.parametrize("side", [256, 1024])
def test_direct_split_variant_matches_tile_kernel_multichannel(O, side):
    """Several kernels per channel: variant 2 (a workgroup per tile and
    kernel, then a pass that adds each channel's growth planes in kernel
    order) forms the same sums as the tile kernel, so the bits agree."""
    d = W.multichannel_1024(seed=4, side=side, path="direct_bf16")
    C = d.channels
    s = np.random.default_rng(side).random((C, side, side), dtype=np.float32)
    got = _run_variants(d, s, 3)
    splits = [v for v in got if v[2] == 2]
    assert splits and [v for v in got if v[2] == 0], sorted(got)
    assert not [v for v in got if v[2] == 1]
    ref = got[(128, 64, 0)] if (128, 64, 0) in got else got[(64, 64, 0)]
    for v, g in got.items():
        assert np.array_equal(g, ref), v
    sim = Sim(d, 1)
    sim.set_direct_tile(*splits[0])
    sim.write_field("state", s)
    sim.step(1)
    err = np.abs(sim.read_field("state") - O.step64(d, s, "fft")).max()
    print(f"multichannel {side}: variants {sorted(got)}; split step max-abs vs FP64 {err:.3e}")
    assert err <= 1e-3
