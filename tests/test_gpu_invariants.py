"""Size-independent properties of the step at the full 4096^2 size, where
the FP64 oracle would take minutes: the world is a torus and the kernel is
the same at every cell, so a step commutes with translations; the potential
is linear in the state.

Direct path: every output sums the same bf16 products in the same MFMA
order whenever a shift keeps the output's position within its 16-column
strip (dx a multiple of 16; any dy), so those shifts commute bit for bit.
FFT path: the FFT's rounding depends on position, so the properties hold to
the FP32 bar (1e-5 on potentials and states of magnitude <= 1).
"""
import numpy as np
import pytest

from auto_amd import Sim, workloads as W

pytestmark = pytest.mark.gpu

SHIFTS = [(37, 1008), (-1025, 16), (2048, -2048)]


def _step_from(d, s, steps=1):
    sim = Sim(d, 1)
    sim.write_field("state", s)
    sim.step(steps)
    return sim.read_field("state")


def _seeded(d, seed):
    return Sim(d, seed).read_field("state")


@pytest.mark.parametrize("dy,dx", SHIFTS)
def test_direct_step_commutes_with_translation_bitwise(dy, dx):
    d = W.rings3_4096(path="direct_bf16")
    s = _seeded(d, 7)
    base = _step_from(d, s, 2)
    moved = _step_from(d, np.roll(s, (dy, dx), axis=(1, 2)), 2)
    assert np.array_equal(moved, np.roll(base, (dy, dx), axis=(1, 2)))
    assert base.any() and not np.array_equal(base, s)


@pytest.mark.parametrize("dy,dx", [(37, 1001), (-1025, 3)])
def test_fft_step_commutes_with_translation(dy, dx):
    d = W.rings3_4096()
    s = _seeded(d, 7)
    base = _step_from(d, s)
    moved = _step_from(d, np.roll(s, (dy, dx), axis=(1, 2)))
    err = np.abs(moved - np.roll(base, (dy, dx), axis=(1, 2))).max()
    print(f"shift ({dy}, {dx}): max-abs {err:.3e}")
    assert err <= 1e-5


def _potential(d, s):
    sim = Sim(d, 1)
    sim.write_field("state", s)
    return sim.read_field("potential0").reshape(d.height, d.width)


def test_fft_potential_is_linear():
    d = W.rings3_4096()
    a, b = _seeded(d, 3), _seeded(d, 4)
    ua, ub, uab = _potential(d, a), _potential(d, b), _potential(d, 0.5 * a + 0.25 * b)
    err = np.abs(uab - (0.5 * ua + 0.25 * ub)).max()
    print(f"linearity max-abs {err:.3e} (potentials up to {ua.max():.3f})")
    assert err <= 1e-5


def test_fft_potential_of_impulse_is_the_kernel_everywhere():
    """The potential of a unit impulse at (y, x) is the kernel centred
    there: the same taps at every position, summing to 1."""
    d = W.rings3_4096()
    k0 = None
    for y, x in [(0, 0), (1234, 4000), (4095, 17)]:
        s = np.zeros((1, d.height, d.width), np.float32)
        s[0, y, x] = 1.0
        u = np.roll(_potential(d, s), (-y, -x), axis=(0, 1))
        if k0 is None:
            k0 = u
            assert abs(float(u.sum(dtype=np.float64)) - 1.0) < 1e-4
        else:
            assert np.abs(u - k0).max() <= 1e-6


def test_smoothlife_4096_step_commutes_with_translation():
    """SmoothLife at the full 4096^2 size (the per-step bar against FP64 at
    smaller sizes is 1e-4; its transition is continuous)."""
    d = W.smoothlife(4096, 4096)
    s = _seeded(d, 5)
    base = _step_from(d, s, 2)
    moved = _step_from(d, np.roll(s, (1111, -333), axis=(1, 2)), 2)
    err = np.abs(moved - np.roll(base, (1111, -333), axis=(1, 2))).max()
    print(f"smoothlife 4096^2 shift: max-abs {err:.3e}, mean {base.mean():.4f}")
    assert err <= 1e-4
    assert 0 < base.mean() < 1


def test_registry_lenia_8192_steps_and_commutes_with_translation():
    """automata::build("lenia", 8192, 8192) (the reference benches Lenia at
    8192^2, results/gpu.jsonl:82): the seeded world steps, stays in [0, 1],
    and a step commutes with a translation across the wrap to the FP32 bar."""
    from auto_amd import build
    d = build("lenia", 8192, 8192)
    assert (d.width, d.height) == (8192, 8192)
    s = _seeded(d, 5)
    assert s.any()
    base = _step_from(d, s, 2)
    assert np.isfinite(base).all() and base.min() >= 0 and base.max() <= 1
    assert not np.array_equal(base, s)
    moved = _step_from(d, np.roll(s, (4111, -2049), axis=(1, 2)), 2)
    err = np.abs(moved - np.roll(base, (4111, -2049), axis=(1, 2))).max()
    print(f"lenia 8192^2 shift: max-abs {err:.3e}, mass {base.sum(dtype=np.float64):.1f}")
    assert err <= 1e-5


def test_fft_potential_of_impulse_is_the_kernel_8192():
    """At 8192^2 the potential of a unit impulse is the same normalised
    kernel wherever the impulse sits."""
    d = W.rings3_4096(side=8192)
    k0 = None
    for y, x in [(0, 0), (5000, 8191), (8191, 3)]:
        s = np.zeros((1, d.height, d.width), np.float32)
        s[0, y, x] = 1.0
        u = np.roll(_potential(d, s), (-y, -x), axis=(0, 1))
        if k0 is None:
            k0 = u
            assert abs(float(u.sum(dtype=np.float64)) - 1.0) < 1e-4
            # The taps lie within the radius: nothing beyond |dx|, |dy| > 26.
            far = np.ones_like(u, bool)
            far[:27, :27] = far[:27, -26:] = far[-26:, :27] = far[-26:, -26:] = False
            assert np.abs(u[far]).max() <= 1e-6
        else:
            assert np.abs(u - k0).max() <= 1e-6
