"""GPU parity of the plain FP32 FFT path against the oracle.

Mirrors the reference's test (crates/auto-core/tests/automata.rs:300-362):
the GPU step from a seeded state must match a plain CPU version of the rule.
Every check here compares the HIP engine (through the C ABI) with the
oracle on the same input; tolerances are written in each test.
"""
import numpy as np
import pytest

from auto_amd import Sim, workloads as W
from auto_amd._abi import INIT_ZERO, EngineError

pytestmark = pytest.mark.gpu


def first_orbium_seed(O, d):
    for seed in range(1, 64):
        if (O.init_state(d, seed) > 0).any():
            return seed
    raise AssertionError("no seed places an Orbium")


def test_reference_known_answer(O):
    """automata.rs:300-362: 128x64, R=13, Orbium; one step vs FP64 direct < 0.002."""
    d = W.lenia_reference(128, 64)
    seed = first_orbium_seed(O, d)
    sim = Sim(d, seed)
    s0 = sim.read_field("state")
    assert np.array_equal(s0, O.init_state(d, seed))
    sim.step(1)
    got = sim.read_field("state")
    want = O.step64(d, s0, "direct")
    err = np.abs(got - want).max()
    print(f"seed {seed} max-abs vs FP64 direct {err:.3e}")
    assert err < 0.002
    assert err < 2e-5  # this engine's own bar: FP64-generated twiddles and kernel spectrum
    assert np.abs(got - O.step_ref32(d, s0)).max() < 2e-5


@pytest.mark.parametrize("w,h", [(16, 16), (32, 64), (64, 32), (128, 128), (256, 64), (512, 1024), (2048, 256),
                                 (64, 4096), (2048, 4096), (8192, 64), (8192, 16), (64, 8192), (256, 8192)])
def test_conv_matches_fp64(O, w, h):
    """The convolution alone (the potential of the current state) vs FP64, random state."""
    radius = min(w, h) / 2 - 1.5
    d = W.desc(w, h, [dict(radius=min(radius, 13.0), beta=[1.0, 0.5], mu=0.2, sigma=0.03)], 0.1,
               init=INIT_ZERO)
    sim = Sim(d, 1)
    rng = np.random.default_rng(w * 7 + h)
    s = rng.random((1, h, w), dtype=np.float32)
    sim.write_field("state", s)
    got = sim.read_field("potential0")
    want = O.conv_fft64(d, s, 0)
    err = np.abs(got - want).max()
    print(f"{w}x{h} conv max-abs {err:.3e}")
    assert err < 1e-5
    # The state is unchanged by reading a conv field.
    assert np.array_equal(sim.read_field("state"), s)


@pytest.mark.parametrize("init,seed", [("orbium", 5), ("patch", 3), ("blobs", 7)])
def test_init_bit_exact(O, init, seed):
    if init == "orbium":
        d = W.lenia_reference(512, 256)
    elif init == "patch":
        d = W.multichannel_1024(seed=2, side=256)
    else:
        d = W.rings3_4096(side=512)
    sim = Sim(d, seed)
    assert np.array_equal(sim.read_field("state"), O.init_state(d, seed))


def test_single_ring_256_100_steps(O):
    """256^2, R=13, 64^2 patch, 100 steps: every step vs the reference FP32 step
    and the FP64 step from the same input state."""
    d = W.single_ring_256()
    sim = Sim(d, 1)
    spectra = O.kernel_spectra32(d)
    khats = O.kernel_hats64(d)
    s = sim.read_field("state")
    worst32 = worst64 = 0.0
    for t in range(100):
        sim.step(1)
        got = sim.read_field("state")
        e32 = np.abs(got - O.step_ref32(d, s, spectra)).max()
        e64 = np.abs(got - O.step64(d, s, "fft", khats)).max()
        worst32, worst64 = max(worst32, e32), max(worst64, e64)
        s = got
    print(f"100 steps: worst per-step max-abs vs FP32 ref {worst32:.3e}, vs FP64 {worst64:.3e}")
    assert worst32 < 1e-4 and worst64 < 1e-4
    assert s.max() > 0  # the patch did not die out numerically


def test_rings3_4096_one_step(O):
    """4096^2 three-ring rule: one step vs FP64 from the same input (<= 1e-3 max-abs)."""
    d = W.rings3_4096()
    sim = Sim(d, 11)
    sim.step(3)
    s = sim.read_field("state")
    sim.step(1)
    got = sim.read_field("state")
    want = O.step64(d, s, "fft")
    diff = got.astype(np.float64) - want
    err = np.abs(diff).max()
    rel = np.linalg.norm(diff) / np.linalg.norm(want)
    print(f"4096^2 rings3 max-abs {err:.3e} rel-L2 {rel:.3e}")
    assert err <= 1e-3
    assert err < 1e-4


def test_multichannel_steps(O):
    """3 channels, 9 kernels (k = 3*src+dst), per-step vs FP64 from the same input."""
    d = W.multichannel_1024(seed=4, side=256)
    sim = Sim(d, 9)
    khats = O.kernel_hats64(d)
    s = sim.read_field("state")
    worst = 0.0
    for _ in range(5):
        sim.step(1)
        got = sim.read_field("state")
        worst = max(worst, np.abs(got - O.step64(d, s, "fft", khats)).max())
        s = got
    print(f"multichannel worst per-step max-abs {worst:.3e}")
    assert worst < 1e-4


def test_multichannel_resume_bitwise():
    """Multichannel rows (several kernels summed per channel) carry the next
    step's spectra too: a run resumed from a saved state equals the
    uninterrupted run."""
    d = W.multichannel_1024(seed=4, side=256)
    a, b = Sim(d, 3), Sim(d, 3)
    a.step(3)
    b.write_field("state", a.read_field("state"))
    a.step(5)
    b.step(5)
    assert np.array_equal(a.read_field("state"), b.read_field("state"))


def test_stats_match_oracle(O):
    d = W.rings3_4096(side=1024)
    sim = Sim(d, 3)
    sim.step(2)
    st = sim.stats()
    s = sim.read_field("state")
    ref = O.stats(s, 1)
    assert abs(st.mass[0] - ref.mass[0]) <= 1e-9 * max(1.0, ref.mass[0])
    assert st.min[0] == ref.min[0] and st.max[0] == ref.max[0]
    assert st.nonfinite == 0 and st.generation == 2


def test_stats_count_nonfinite():
    d = W.single_ring_256()
    sim = Sim(d, 1)
    s = sim.read_field("state")
    s[0, 3, 5] = np.nan
    s[0, 7, 9] = np.inf
    sim.write_field("state", s)
    st = sim.stats()
    assert st.nonfinite == 2
    assert st.mass[0] == pytest.approx(float(np.nansum(np.where(np.isfinite(s), s, 0), dtype=np.float64)))


def test_check_reports_a_diverged_state():
    """ah_sim_check: AH_OK on a finite state, AH_ERR_DIVERGED naming the
    count once a NaN is written in.  A step from that state spreads the NaN
    through the FFT to every potential, and the clip (fmax/fmin return the
    non-NaN operand) maps every cell to 0: the step never leaves a non-finite
    state behind, so the check belongs before stepping a state from outside."""
    from auto_amd._abi import EngineError, ERR_DIVERGED
    d = W.single_ring_256()
    sim = Sim(d, 1)
    sim.step(3)
    sim.check()
    s = sim.read_field("state")
    clean = s.copy()
    s[0, 100, 100] = np.nan
    sim.write_field("state", s)
    with pytest.raises(EngineError) as e:
        sim.check()
    assert e.value.code == ERR_DIVERGED and "1 non-finite" in str(e.value)
    sim.step(1)
    sim.check()
    after = sim.read_field("state")
    assert np.isfinite(after).all() and after.max() == 0.0
    sim.write_field("state", clean)
    sim.check()


def test_deterministic_and_resumable():
    """Same seed -> bitwise-equal states; save at t and continue == uninterrupted run."""
    d = W.rings3_4096(side=512)
    a, b = Sim(d, 21), Sim(d, 21)
    a.step(10)
    b.step(4)
    saved = b.read_field("state")
    c = Sim(d, 999)
    c.step(2)  # c now holds spectra of its own state, which the write must drop
    c.write_field("state", saved)
    c.step(6)
    b.step(6)
    sa = a.read_field("state")
    assert np.array_equal(sa, b.read_field("state"))
    assert np.array_equal(sa, c.read_field("state"))


def test_steps_split_across_calls_and_conv_reads():
    """A step reuses the spectra the previous step's last pass left; one
    10-step call, ten 1-step calls and 1-step calls with potential reads in
    between (which consume those spectra) give bitwise-equal states."""
    d = W.rings3_4096(side=512)
    a, b, c = Sim(d, 4), Sim(d, 4), Sim(d, 4)
    a.step(10)
    for _ in range(10):
        b.step(1)
    for t in range(10):
        c.step(1)
        if t % 3 == 1:
            c.read_field("potential0")
            c.read_field("conv0")
    sa = a.read_field("state")
    assert np.array_equal(sa, b.read_field("state"))
    assert np.array_equal(sa, c.read_field("state"))
    assert a.info().generation == c.info().generation == 10


def test_field_errors():
    sim = Sim(W.single_ring_256(), 1)
    with pytest.raises(EngineError):
        sim.read_field("nope")
    with pytest.raises(EngineError):
        sim.read_field("conv3")
    with pytest.raises(EngineError):
        sim.read_field("potential3")
    buf = np.zeros(4, np.float32)
    assert sim._lib.ah_sim_read_field(sim._h, b"state", buf.ctypes.data, buf.nbytes) == 5  # AH_ERR_SIZE
    assert b"bytes" in sim._lib.ah_last_error()


def test_profile_counts_sampled_launches():
    """profile(every=N) times every N-th launch of each kernel; a negative
    setting is rejected with the C error code."""
    from auto_amd._abi import ERR_INVALID
    sim = Sim(W.rings3_4096(side=512), 2)
    sim.profile(True, every=4)
    sim.step(40)
    p = sim.profile_read()
    assert p["col_conv"][0] == 10 and p["row_c2r_update_r2c"][0] == 10
    assert p["row_r2c"][0] == 1  # only the first step transforms the state itself
    assert all(ms > 0 for _, ms in p.values())
    sim.profile(True)
    sim.step(3)
    assert sim.profile_read()["col_conv"][0] == 3
    assert sim._lib.ah_sim_profile(sim._h, -1) == ERR_INVALID


@pytest.mark.parametrize("w,h", [(4096, 64), (4096, 256), (8192, 128), (2048, 64), (1024, 256), (512, 128),
                                 (256, 512), (128, 128), (128, 16), (64, 8192), (128, 8192)])
def test_wide_world_steps_and_resume(O, w, h):
    """Worlds 128..4096 wide run the fused row pass (2048/M rows per
    workgroup), 8192-wide ones (the widest a descriptor allows) and worlds
    with fewer rows than a fused workgroup takes (128x16) the generic one:
    every step vs FP64 from the same input, and a run resumed from a saved
    state (whose first step transforms the state in pass 1) is bitwise equal
    to the uninterrupted run (whose spectra come from the previous step's
    pass 3)."""
    k = dict(W.RINGS3, radius=float(min(26, min(w, h) // 2 - 2)))
    d = W.desc(w, h, [k], 0.1, init=W.INIT_BLOBS, init_size=min(48, h // 2), init_count=16)
    a = Sim(d, 5)
    khats = O.kernel_hats64(d)
    s = a.read_field("state")
    worst = 0.0
    for _ in range(4):
        a.step(1)
        got = a.read_field("state")
        worst = max(worst, np.abs(got - O.step64(d, s, "fft", khats)).max())
        s = got
    print(f"{w}x{h} worst per-step max-abs {worst:.3e}")
    assert worst < 1e-5
    b = Sim(d, 77)
    b.write_field("state", s)
    a.step(6)
    b.step(6)
    assert np.array_equal(a.read_field("state"), b.read_field("state"))


@pytest.mark.parametrize("path", ["fft_fp32", "direct_bf16"])
def test_conv_field_is_the_last_steps_potential(path):
    """conv<k> follows the reference's conv_i (spectral.rs:227-234): zero
    before the first step, then the potential the last step used (that of
    its input state), unchanged by a state write; potential<k> is the
    current state's.  Both are the same computation on the same input, so
    they agree bit for bit."""
    d = W.rings3_4096(side=256, path=path)
    sim = Sim(d, 3)
    assert not sim.read_field("conv0").any()
    p_before = sim.read_field("potential0")
    assert p_before.any()
    sim.step(1)
    assert np.array_equal(sim.read_field("conv0"), p_before)
    s1 = sim.read_field("state")
    p1 = sim.read_field("potential0")
    sim.write_field("state", np.zeros_like(s1))
    assert np.array_equal(sim.read_field("conv0"), p_before)
    sim.write_field("state", s1)
    sim.step(1)
    assert np.array_equal(sim.read_field("conv0"), p1)
    assert not np.array_equal(p1, p_before)
