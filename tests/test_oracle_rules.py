"""Oracle pinning for the SmoothLife and MNCA rules (CPU).

The reference defines both automata in crates/auto-core/src/automata/
spectral.rs:354-480 and stores no vectors for them, so each oracle function
is checked against something its formula does not restate: integer lattice
counts, areas, discriminating closed-form cases of the transitions, and the
oracle's two independent steps (reference FP32 passes vs FP64 FFT).
"""
import math

import numpy as np
import pytest

from auto_amd import workloads as W
from auto_amd._abi import Kernel, SHAPE_SMOOTH_DISC, SHAPE_SMOOTH_RING, SHAPE_BAND


def kern(shape, radius, inner=0.0, aa=0.0):
    k = Kernel()
    k.shape, k.radius, k.inner, k.aa = shape, radius, inner, aa
    return k


def raw_box(O, k):
    taps = O.kernel_box64(k)
    return taps, O.lib().or_kernel_box64(k, O._dp(np.empty_like(taps)))


def test_band_sum_is_the_integer_lattice_count(O):
    """MNCA ring [lo, hi): the unnormalised box sum counts the lattice points
    with lo^2 <= x^2 + y^2 < hi^2, counted here in integers."""
    for lo, hi in [(12, 18), (3, 8), (0, 5), (10, 15)]:
        _, raw = raw_box(O, kern(SHAPE_BAND, float(hi), float(lo)))
        n = sum(1 for x in range(-hi, hi + 1) for y in range(-hi, hi + 1) if lo * lo <= x * x + y * y < hi * hi)
        assert raw == n, (lo, hi)


def test_smooth_shapes_closed_forms(O):
    """Disc: 1 inside ri-aa, 1/2 at ri, 0 from ri+aa; sum -> area pi ri^2.
    Ring: 1/2 at both edges, 1 mid-ring, 0 inside ri-aa and beyond ra+aa;
    sum -> annulus area pi (ra^2 - ri^2)."""
    ri, ra, aa = 7.0, 21.0, 0.5
    disc, ring = kern(SHAPE_SMOOTH_DISC, ri, aa=aa), kern(SHAPE_SMOOTH_RING, ra, ri, aa)
    p = lambda k, x, y=0: O.lib().or_profile64(k, x, y)
    assert p(disc, 0) == 1.0 and p(disc, 6) == 1.0 and p(disc, 7) == 0.5 and p(disc, 8) == 0.0
    assert p(ring, 7) == 0.5 and p(ring, 21) == 0.5 and p(ring, 14) == 1.0
    assert p(ring, 6) == 0.0 and p(ring, 22) == 0.0 and p(ring, 0) == 0.0
    assert p(ring, 3, 4) == 0.0  # r = 5 exactly
    _, sd = raw_box(O, disc)
    _, sr = raw_box(O, ring)
    assert abs(sd - math.pi * ri * ri) / (math.pi * ri * ri) < 0.01
    assert abs(sr - math.pi * (ra * ra - ri * ri)) / (math.pi * (ra * ra - ri * ri)) < 0.005
    assert O.lib().or_kernel_reach(disc) == 8 and O.lib().or_kernel_reach(ring) == 22


def test_smoothlife_transition_cases(O):
    """Inner average 0 selects the birth interval [B1, B2] = [.278, .365];
    inner average 1 the survival interval [D1, D2] = [.267, .445].  A ring
    average of .40 is outside birth but inside survival, which a swapped mix
    or swapped interval would get wrong; far outside both, the cell decays."""
    f64, f32 = O.lib().or_smoothlife64, O.lib().or_smoothlife32
    dt = 0.1
    assert f64(0.5, 0.0, 0.32, dt) == pytest.approx(0.5 + dt, abs=2e-3)   # birth
    assert f64(0.5, 0.0, 0.40, dt) == pytest.approx(0.5 - dt, abs=2e-3)   # no birth at .40
    assert f64(0.5, 1.0, 0.40, dt) == pytest.approx(0.5 + dt, abs=2e-3)   # survival at .40
    assert f64(0.5, 1.0, 0.20, dt) == pytest.approx(0.5 - dt, abs=2e-3)   # dies below D1
    assert f64(0.05, 0.0, 0.9, dt) == 0.0 and f64(0.95, 0.0, 0.32, dt) == 1.0  # clip
    rng = np.random.default_rng(3)
    for s, c0, c1 in rng.random((200, 3)):
        assert abs(f32(s, c0, c1, dt) - f64(s, c0, c1, dt)) < 2e-5


def test_mnca_transition_cases(O):
    """Later bands overrule earlier ones; bands are closed intervals on the
    f32 thresholds."""
    f64 = O.lib().or_mnca64
    f32 = O.lib().or_mnca32
    assert f64(0.0, 0.215, 0.0) == 1.0          # birth band of u0
    assert f64(0.0, 0.215, 0.2) == 0.0          # u1 death band overrules
    assert f64(1.0, 0.4, 0.0) == 0.0            # u0 death band
    assert f64(1.0, 0.4, 0.5) == 1.0            # u1 birth band overrules
    assert f64(0.7, 0.3, 0.3) == 0.7            # no band: state kept
    lo = float(np.float32(0.210))
    hi = float(np.float32(0.220))
    assert f64(0.0, lo, 0.0) == 1.0 and f64(0.0, hi, 0.0) == 1.0
    assert f64(0.0, float(np.nextafter(np.float32(hi), np.float32(1))), 0.0) == 0.0
    rng = np.random.default_rng(4)
    for s, c0, c1 in rng.random((500, 3)).astype(np.float32):
        assert f32(s, c0, c1) == f64(float(s), float(c0), float(c1))


def test_disc_and_noise_seeds(O):
    """SmoothLife seed: every tile is empty or holds exactly the lattice disc
    (a, b) in [-48, 48)^2 with a^2 + b^2 < 576, centred at (48, 48).
    MNCA seed: 0/1 values, constant on 3x3 blocks, about 32% alive."""
    d = W.smoothlife(384, 384)
    s = O.init_state(d, 11)[0]
    disc = np.array([[1.0 if a * a + b * b < 576 else 0.0 for a in range(-48, 48)] for b in range(-48, 48)])
    full = 0
    for ty in range(4):
        for tx in range(4):
            t = s[96 * ty:96 * ty + 96, 96 * tx:96 * tx + 96]
            assert t.sum() == 0 or np.array_equal(t, disc)
            full += t.sum() > 0
    assert 0 < full < 16
    d = W.mnca(258, 258)  # rounds to 512
    n = O.init_state(d, 5)[0]
    assert set(np.unique(n)) <= {0.0, 1.0}
    b = n[:510, :510].reshape(170, 3, 170, 3)
    assert np.all(b == b[:, :1, :, :1])
    assert abs(n.mean() - 0.32) < 0.01


@pytest.mark.parametrize("rule", ["smoothlife", "mnca"])
def test_reference_fp32_step_agrees_with_fp64(O, rule):
    """Two independent oracle steps: the reference's FP32 Stockham passes with
    the f32 transition, and numpy's FP64 FFT with the FP64 transition.  For
    MNCA the two may differ only where an FP64 potential lies within 1e-5 of
    a band edge."""
    d = W.smoothlife(128, 128) if rule == "smoothlife" else W.mnca(128, 128)
    s = O.init_state(d, 2)
    for _ in range(3):  # a few steps so the state is not just the seed
        s = O.step64(d, s, "fft")
    a = O.step_ref32(d, s)
    b = O.step64(d, s, "fft")
    if rule == "smoothlife":
        assert np.abs(a - b).max() < 1e-4
        assert np.abs(b - s).max() > 0.05  # the step does something
        return
    u = O.potentials64(d, s)
    edges0 = np.float32([0.210, 0.220, 0.350, 0.500, 0.750, 0.850, 0.120, 0.150])
    edges1 = np.float32([0.100, 0.280, 0.430, 0.550])
    near = np.zeros(u.shape[1:], bool)
    for e in edges0:
        near |= np.abs(u[0] - e) < 1e-5
    for e in edges1:
        near |= np.abs(u[1] - e) < 1e-5
    diff = (a != b)[0]
    assert not np.any(diff & ~near)
    assert np.any(a != s)
