"""Batched uneven worlds on the GPU: every world of the batch matches the
plain single-world step, and the graph-replayed loop equals eager stepping."""
import numpy as np
import pytest

from auto_amd import Sim, Batch, workloads as W
from auto_amd._abi import EngineError, LeniaDesc, PATH_FFT_FP32

pytestmark = pytest.mark.gpu


def small_worlds(n=24, seed=3):
    return W.uneven_worlds(seed=seed, n=n, path="direct_bf16", sides=(128, 256, 512))


def test_batch_worlds_match_single_world_step(O):
    """Per world: batch step vs the plain FP32 FFT single-world step and vs FP64,
    from the same input state (<= 1e-3 max-abs)."""
    worlds = small_worlds()
    b = Batch(worlds)
    info = b.info()
    assert info.n_worlds == len(worlds) and info.n_tune >= 1
    b.step(2)
    before = [b.read_world(i) for i in range(len(worlds))]
    b.step(1)
    worst_plain = worst64 = 0.0
    for i, (d, seed) in enumerate(worlds):
        got = b.read_world(i)
        plain = LeniaDesc.from_buffer_copy(d)
        plain.path = PATH_FFT_FP32
        sim = Sim(plain, seed)
        sim.write_field("state", before[i][None])
        sim.step(1)
        ref = sim.read_field("state")[0]
        worst_plain = max(worst_plain, float(np.abs(got - ref).max()))
        worst64 = max(worst64, float(np.abs(got - O.step64(d, before[i][None], "fft")[0]).max()))
    print(f"batch of {len(worlds)}: worst max-abs vs plain FFT step {worst_plain:.3e}, vs FP64 {worst64:.3e}; "
          f"tiles {info.ntiles} shape {info.tile_cols}x{info.tile_rows} grid {info.grid}")
    assert worst_plain <= 1e-3 and worst64 <= 1e-3


def test_batch_seeds_match_single_sims(O):
    worlds = small_worlds(n=6, seed=8)
    b = Batch(worlds)
    for i, (d, seed) in enumerate(worlds):
        assert np.array_equal(b.read_world(i), O.init_state(d, seed)[0])


def test_batch_graph_equals_eager():
    """100 steps through one graph replay == 100 eager steps (bitwise)."""
    worlds = small_worlds(n=8, seed=5)
    a, b = Batch(worlds), Batch(worlds)
    a.step(100)          # one graph replay
    for _ in range(100):  # 100 single launches
        b.step(1)
    for i in range(len(worlds)):
        assert np.array_equal(a.read_world(i), b.read_world(i))
    # Replays after the first give the same as eager too.
    a.step(100)
    for _ in range(100):
        b.step(1)
    for i in range(len(worlds)):
        assert np.array_equal(a.read_world(i), b.read_world(i))


def test_batch_write_world_roundtrip():
    worlds = small_worlds(n=4, seed=9)
    b = Batch(worlds)
    d = worlds[2][0]
    x = np.random.default_rng(0).random((d.height, d.width), dtype=np.float32)
    b.write_world(2, x)
    assert np.array_equal(b.read_world(2), x)
    buf = np.zeros(16, np.float32)
    assert b._lib.ah_batch_read_world(b._h, 99, buf.ctypes.data, buf.nbytes) == 1  # AH_ERR_INVALID
    assert b._lib.ah_batch_read_world(b._h, 1, buf.ctypes.data, buf.nbytes) == 5  # AH_ERR_SIZE


def test_batch_measure_reports_overhead():
    b = Batch(small_worlds(n=16, seed=2))
    m = b.measure(100)
    assert m["kernel_ms"] > 0 and m["graph_wall_ms"] > 0
    print(m)


def test_batch_rejects_fft_path():
    d, s = W.uneven_worlds(seed=1, n=1, path="fft_fp32")[0]
    with pytest.raises(EngineError):
        Batch([(d, s)])
