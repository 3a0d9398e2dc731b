"""CPU checks that pin the oracle itself (no GPU).

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign, index or normalisation) fails it: published generator outputs,
brute-force DFTs, an independent FP64 FFT (numpy), the reference's own
known-answer tolerance (crates/auto-core/tests/automata.rs:300-362), closed
forms of the kernel and the growth map, and the committed golden fixtures.
"""
import os

import numpy as np
import pytest

from auto_amd import workloads as W
from auto_amd._abi import INIT_ZERO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------ generators

def test_splitmix64_published_vector(O):
    """SplitMix64 from state 0: its published first outputs."""
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert O.splitmix64(0, 3) == want
    r = W.SplitMix64(0)
    assert [r.next() for _ in range(3)] == want


def test_pcg_rand3_unit_f_golden_vectors(O):
    """pcg / rand3 / unit_f against vectors evaluated from the reference's WGSL
    text (program.rs:139-146; constants parsed, not retyped: see
    tests/golden/make_pcg_vectors.py)."""
    g = np.load(os.path.join(GOLDEN, "pcg_vectors.npz"))
    assert [O.pcg(int(v)) for v in g["v"]] == [int(x) for x in g["pcg"]]
    got = [O.rand3(int(a), int(b), int(c)) for a, b, c in zip(g["a"], g["b"], g["c"])]
    assert got == [int(x) for x in g["rand3"]]
    got = np.array([O.unit_f(int(r)) for r in g["r"]], np.float32)
    assert np.array_equal(got.view(np.uint32), g["unit_f"].view(np.uint32))


def test_pcg_is_a_bijection_on_a_sample(O):
    """Every step of the PCG hash is invertible on u32, so distinct inputs give distinct outputs."""
    outs = {O.pcg(v) for v in range(0, 1 << 16)}
    outs |= {O.pcg(v) for v in range((1 << 31) - 4096, (1 << 31) + 4096)}
    assert len(outs) == (1 << 16) + 8192


def test_unit_f_range(O):
    vals = [O.unit_f(O.rand3(x, 7, 3)) for x in range(5000)]
    assert 0.0 <= min(vals) and max(vals) < 1.0
    assert abs(np.mean(vals) - 0.5) < 0.02
    assert O.unit_f(0) == 0.0 and O.unit_f(0xFFFFFFFF) == 1.0  # f32 rounding of (2^32-1)/2^32


def test_orbium_kat_places_a_creature(O):
    """automata.rs:305-314: some seed in 1..63 places an Orbium at 128x64."""
    d = W.lenia_reference(128, 64)
    seeds = [s for s in range(1, 64) if (O.init_state(d, s) > 0).any()]
    assert seeds
    s = O.init_state(d, seeds[0])[0]
    # Each placed creature is a quarter-turn of the same 20x20 tile: equal sums.
    tiles = [s[ty:ty + 20, tx:tx + 20] for ty in range(0, 64, 64) for tx in range(0, 128, 64)]
    sums = {round(float(t.sum()), 4) for t in tiles if t.any()}
    assert len(sums) == 1
    assert s[(np.arange(64) % 64 >= 20)].max() == 0.0  # nothing outside the 20-cell corner


def test_blob_init_semantics(O):
    d = W.rings3_4096(side=256)
    s = O.init_state(d, 5)[0]
    st = W.SplitMix64(5)
    mask = np.zeros((256, 256), bool)
    for _ in range(64):
        x0, y0 = st.next() % 256, st.next() % 256
        ys = (np.arange(48) + y0) % 256
        xs = (np.arange(48) + x0) % 256
        mask[np.ix_(ys, xs)] = True
    assert not s[~mask].any()
    y, x = np.argwhere(mask)[123]
    assert s[y, x] == O.unit_f(O.rand3(int(x), int(y), 5))


# ------------------------------------------------------------------ FFT

@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64])
def test_stockham_matches_bruteforce_dft(O, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n)
    want = x.astype(np.complex128) @ F.T
    got = O.fft_axis_ref32(x, True, False)
    assert np.abs(got - want).max() < 1e-5 * n
    # inverse: sign +1, unnormalised (spectral.rs:225-226)
    back = O.fft_axis_ref32(got.astype(np.complex64), True, True)
    assert np.abs(back - n * x).max() < 1e-4 * n


def test_stockham_columns_match_numpy(O):
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((256, 32)) + 1j * rng.standard_normal((256, 32))).astype(np.complex64)
    got = O.fft_axis_ref32(x, False, False)
    assert np.abs(got - np.fft.fft(x.astype(np.complex128), axis=0)).max() < 1e-3


# --------------------------------------------------------------- kernels

def test_core_closed_form(O):
    assert O.lib().or_core64(0.5) == 1.0  # exp(4 - 1/(1/4)) = exp(0)
    assert O.lib().or_core64(0.0) == 0.0 and O.lib().or_core64(1.0) == 0.0
    q = 0.3
    assert O.lib().or_core64(q) == pytest.approx(np.exp(4 - 1 / (q * (1 - q))), rel=1e-15)


def test_multiring_shell_peaks(O):
    """Ring k of a 3-ring shell peaks at r = R(k+1/2)/3 with height beta_k."""
    import ctypes
    beta = (ctypes.c_float * 3)(1.0, 2.0 / 3.0, 1.0 / 3.0)
    R = 26.0
    for k in range(3):
        r = R * (k + 0.5) / 3
        assert O.lib().or_shell64(r, R, 3, beta) == pytest.approx(beta[k], rel=1e-12)
    assert O.lib().or_shell64(R, R, 3, beta) == 0.0
    assert O.lib().or_shell64(R * 2 / 3, R, 3, beta) == 0.0  # ring boundary


@pytest.mark.parametrize("radius,beta", [(13.0, [1.0]), (26.0, [1.0, 2 / 3, 1 / 3]), (9.5, [0.3, 1.0])])
def test_kernel_taps_normalised_and_symmetric(O, radius, beta):
    d = W.desc(128, 128, [dict(radius=radius, beta=beta, mu=0.2, sigma=0.02)], 0.1)
    t = O.kernel_box64(d.kernels[0])
    assert t.sum() == pytest.approx(1.0, abs=1e-13)
    assert np.array_equal(t, t[::-1, :]) and np.array_equal(t, t[:, ::-1]) and np.array_equal(t, t.T)
    reach = t.shape[0] // 2
    yy, xx = np.mgrid[-reach:reach + 1, -reach:reach + 1]
    assert not t[np.hypot(xx, yy) >= radius].any()


def test_kernel_matches_reference_normalisation(O):
    """The FP64 kernel spectrum agrees with the reference init's FP32 spectrum
    (f32 shape, box sum over ceil(R)+1, spectral.rs:75-83,150-181)."""
    d = W.lenia_reference(64, 64)
    ref = O.kernel_spectrum_ref32(64, 64, 13.0)
    k64 = np.fft.fft2(O.kernel_torus64(d.kernels[0], 64, 64))
    assert np.abs(ref - k64).max() < 2e-6
    assert np.abs(k64.imag).max() < 1e-15  # even kernel -> real spectrum


# ---------------------------------------------------------------- steps

def kat_state(O):
    d = W.lenia_reference(128, 64)
    seed = next(s for s in range(1, 64) if (O.init_state(d, s) > 0).any())
    return d, seed, O.init_state(d, seed)


def test_reference_known_answer_fp32_step(O):
    """The restated FP32 step passes the reference's own test: < 0.002 vs FP64 direct."""
    d, _, s0 = kat_state(O)
    got = O.step_ref32(d, s0)
    want = O.step64(d, s0, "direct")
    assert np.abs(got - want).max() < 0.002
    assert np.abs(got - want).max() < 1e-5


@pytest.mark.parametrize("which", ["rings3", "multichannel"])
def test_direct64_equals_fft64(O, which):
    """Two independent FP64 convolutions agree (brute-force sum vs numpy's FP64 FFT)."""
    if which == "rings3":
        d = W.desc(128, 64, [dict(W.RINGS3, radius=20.0)], 0.1, init=INIT_ZERO)
    else:
        d = W.multichannel_1024(seed=3, side=64)
    rng = np.random.default_rng(0)
    s = rng.random((d.channels, d.height, d.width), dtype=np.float32)
    for k in range(d.n_kernels):
        assert np.abs(O.conv_direct64(d, s, k) - O.conv_fft64(d, s, k)).max() < 1e-12
    assert np.abs(O.step64(d, s, "direct") - O.step64(d, s, "fft")).max() <= 6e-8


def test_uniform_field_is_a_fixed_point_of_the_convolution(O):
    d = W.multichannel_1024(seed=5, side=64)
    s = np.full((3, 64, 64), 0.37, np.float32)
    for k in range(d.n_kernels):
        assert np.abs(O.conv_direct64(d, s, k) - np.float32(0.37)).max() < 1e-12


def test_growth_closed_form(O):
    d = W.single_ring_256()
    s = np.full((1, 256, 256), 0.5, np.float32)
    u = np.full((1, 256, 256), 0.15)  # u = mu -> growth +1
    assert np.all(O.update64(d, u, s) == np.float32(0.6))
    u[:] = 10.0  # far from mu -> growth -1
    assert np.all(O.update64(d, u, s) == np.float32(0.4))
    s[:] = 0.95
    u[:] = 0.15
    assert np.all(O.update64(d, u, s) == 1.0)  # clipped


def test_multichannel_update_weights(O):
    """s'_c = clip(s_c + dt * sum_{k->c} h_k g_k, 0, 1) with h normalised per destination."""
    d = W.multichannel_1024(seed=1, side=64)
    s = np.full((3, 64, 64), 0.5, np.float32)
    u = np.zeros((9, 64, 64))
    for k in range(9):
        u[k] = d.kernels[k].mu  # every growth is +1 -> weighted sum is exactly 1
    out = O.update64(d, u, s)
    assert np.abs(out - np.float32(0.5 + d.dt)).max() < 1e-7


def test_stats_oracle(O):
    s = np.array([[0.0, 0.5, np.nan, 1.0]], np.float32).reshape(1, 1, 4)
    st = O.stats(s, 1)
    assert st.mass[0] == 1.5 and st.min[0] == 0.0 and st.max[0] == 1.0 and st.nonfinite == 1


# --------------------------------------------------------------- golden

def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_golden_kat(O):
    g = golden("kat_lenia_128x64.npz")
    d, seed, s0 = kat_state(O)
    assert int(g["seed"]) == seed
    assert np.array_equal(g["s0"], s0[0])
    assert np.abs(O.step64(d, s0, "direct")[0] - g["step64"]).max() == 0.0
    assert np.abs(O.step_ref32(d, s0)[0] - g["step_ref32"]).max() < 1e-6


def test_golden_rings3(O):
    g = golden("rings3_128.npz")
    d = W.rings3_4096(side=128)
    d.kernels[0].radius = 20.0
    s0 = O.init_state(d, int(g["seed"]))
    assert np.array_equal(s0, g["s0"])
    assert np.abs(O.step64(d, s0, "direct") - g["step64"]).max() == 0.0


def test_golden_multichannel(O):
    g = golden("multichannel_64.npz")
    d = W.multichannel_1024(seed=int(g["rule_seed"]), side=64)
    s0 = O.init_state(d, int(g["seed"]))
    assert np.array_equal(s0, g["s0"])
    assert np.abs(O.step64(d, s0, "direct") - g["step64"]).max() == 0.0


def test_golden_dft(O):
    g = golden("dft_32.npz")
    got = O.fft_axis_ref32(g["x"], True, False)
    assert np.abs(got - g["X"]).max() < 1e-4
