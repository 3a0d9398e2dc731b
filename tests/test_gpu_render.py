"""Sim.render: the view pass of sim.rs:17-41 with the rules' palettes (GPU).

Checked on states whose picture is known in closed form: uniform fields give
the palette's end stops, a half value gives the mid stop, samples outside
the world are black, and a 2x2 supersampled checkerboard averages in light,
sqrt((a^2 + b^2) / 2) per channel.
"""
import numpy as np
import pytest

from auto_amd import Sim, workloads as W
from auto_amd._abi import ERR_SIZE

pytestmark = pytest.mark.gpu

LENIA = np.float32([[0.01, 0.01, 0.05], [0.08, 0.12, 0.45], [0.1, 0.6, 0.7], [0.95, 0.85, 0.4], [1.0, 0.98, 0.9]])
MNCA = np.float32([[0.02, 0.02, 0.07], [0.75, 0.95, 0.85]])


def unorm(v):
    v = np.clip(np.float32(v), np.float32(0), np.float32(1))
    return np.floor(v * np.float32(255) + np.float32(0.5)).astype(np.uint8)


def sim_with(desc, value):
    sim = Sim(desc, 1)
    sim.write_field("state", np.full((1, desc.height, desc.width), value, np.float32))
    return sim


def test_uniform_states_give_palette_stops():
    d = W.lenia_reference(256, 128)
    for value, stop in ((0.0, 0), (1.0, 4), (0.5, 2), (0.25, 1)):
        img = sim_with(d, value).render(64, 32)
        assert img.shape == (32, 64, 4)
        # within one unorm step: the ramp forms c_k + (c_k+1 - c_k) t in float32
        assert np.all(np.abs(img[..., :3].astype(int) - unorm(LENIA[stop]).astype(int)) <= 1), value
        assert np.all(img[..., 3] == 255)
    img = sim_with(W.lenia_reference(256, 128), 0.0).render(32, 16)
    assert np.all(img[..., :3] == unorm(LENIA[0]))  # t = 0 is exact
    img = sim_with(W.mnca(256, 256), 0.5).render(32, 32)
    assert np.all(np.abs(img[..., :3].astype(int) - unorm((MNCA[0] + MNCA[1]) * np.float32(0.5)).astype(int)) <= 1)


def test_outside_the_world_is_black_and_supersampling_averages_in_light():
    d = W.lenia_reference(64, 64)
    sim = Sim(d, 1)
    yy, xx = np.mgrid[0:64, 0:64]
    sim.write_field("state", ((xx + yy) % 2).astype(np.float32)[None])
    img = sim.render(32, 32, center=(32.0, 32.0), scale=2.0)  # 2x2 samples per pixel
    want = unorm(np.sqrt((LENIA[0] ** 2 + LENIA[4] ** 2) / np.float32(2)))
    assert np.all(np.abs(img[..., :3].astype(int) - want.astype(int)) <= 1)
    img = sim.render(32, 32, center=(0.0, 32.0), scale=2.0)  # left half of the frame lies at x < 0
    assert np.all(img[:, :16, :3] == 0) and np.all(img[:, 16:, :3] > 0)


def test_render_errors_and_real_world():
    sim = Sim(W.smoothlife(512, 256), 3)
    sim.step(5)
    img = sim.render(256, 128)
    assert img[..., :3].std() > 0
    out = np.empty(10, np.uint8)
    assert sim._lib.ah_sim_render(sim._h, 4, 4, 0.0, 0.0, 1.0, out.ctypes.data, out.nbytes) == ERR_SIZE
