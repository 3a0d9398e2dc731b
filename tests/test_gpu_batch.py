"""Batched uneven worlds on the GPU: every world of the batch matches the
plain single-world step, and the graph-replayed loop equals eager stepping."""
import numpy as np
import pytest

from auto_amd import Sim, Batch, workloads as W
from auto_amd._abi import EngineError, LeniaDesc, PATH_DIRECT_BF16, PATH_FFT_FP32

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


def _world(side, radius, sigma, dt, path, rings=(1.0,)):
    return W.desc(side, side, [dict(radius=float(radius), beta=list(rings), mu=0.2, sigma=sigma)], dt, path,
                  init=W.INIT_PATCH, init_size=side // 4)


def test_batch_fft_and_auto_worlds(O):
    """FFT worlds (grouped by size, one world per channel) and AUTO worlds in
    one batch with direct worlds.  An FFT world matches a standalone Sim of
    the same world to FP32 rounding (the batch's column pass is its own fused
    kernel, so the summation order differs): within 1e-5, against ~1e-6 of
    either vs FP64.  A direct world equals a standalone direct Sim bitwise.
    Every world is within 1e-3 of the FP64 step from the same input; AUTO
    takes FFT for a wide, high-gain kernel (3 bf16 terms, ~100 chunks) and
    direct for a small, low-gain one."""
    worlds = W.uneven_worlds(seed=4, n=5, path="fft_fp32", sides=(128, 256))
    worlds += W.uneven_worlds(seed=5, n=5, path="direct_bf16", sides=(128, 256))
    worlds += [(_world(256, 24, 0.01, 0.2, "auto", (1.0, 0.6, 0.3)), 11),   # -> FFT
               (_world(128, 8, 0.05, 0.05, "auto"), 12)]                   # -> direct
    b = Batch(worlds)
    info = b.info()
    sizes = {(d.width, d.height) for d, _ in worlds[:5]} | {(256, 256)}
    assert info.fft_worlds == 6 and info.fft_groups == len(sizes)
    b.step(2)
    before = [b.read_world(i) for i in range(len(worlds))]
    b.step(1)
    worst64 = worst_sim = 0.0
    for i, (d, seed) in enumerate(worlds):
        got = b.read_world(i)
        worst64 = max(worst64, float(np.abs(got - O.step64(d, before[i][None], "fft")[0]).max()))
        if i < 5 or i == 10:
            plain = LeniaDesc.from_buffer_copy(d)
            plain.path = PATH_FFT_FP32
            sim = Sim(plain, seed)
            sim.write_field("state", before[i][None])
            sim.step(1)
            worst_sim = max(worst_sim, float(np.abs(got - sim.read_field("state")[0]).max()))
        else:
            # A direct world equals a standalone direct Sim: the same MFMA
            # order per cell and the same epilogue arithmetic.
            plain = LeniaDesc.from_buffer_copy(d)
            plain.path = PATH_DIRECT_BF16
            sim = Sim(plain, seed)
            sim.write_field("state", before[i][None])
            sim.step(1)
            assert np.array_equal(got, sim.read_field("state")[0]), i
    print(f"mixed batch: {info.fft_worlds} FFT worlds in {info.fft_groups} groups; worst max-abs vs FP64 {worst64:.3e}, "
          f"FFT worlds vs standalone Sims {worst_sim:.3e}")
    assert worst64 <= 1e-3 and worst_sim < 1e-5
    # A write into an FFT world replaces the spectra its next step uses.
    x = np.random.default_rng(1).random((worlds[0][0].height, worlds[0][0].width), dtype=np.float32)
    b.write_world(0, x)
    b.step(1)
    plain = LeniaDesc.from_buffer_copy(worlds[0][0])
    sim = Sim(plain, 1)
    sim.write_field("state", x[None])
    sim.step(1)
    assert np.abs(b.read_world(0) - sim.read_field("state")[0]).max() < 1e-5


def test_batch_of_fft_worlds_only_and_graph():
    """No direct worlds: no tuner candidates, the FFT groups alone, and graph
    replays equal eager steps."""
    worlds = W.uneven_worlds(seed=7, n=4, path="fft_fp32", sides=(128,))
    a, b = Batch(worlds), Batch(worlds)
    assert a.info().n_tune == 0 and a.info().fft_groups == 1 and a.info().ntiles == 0
    a.step(100)
    for _ in range(100):
        b.step(1)
    for i in range(len(worlds)):
        assert np.array_equal(a.read_world(i), b.read_world(i))


def _world_wh(w, h, radius, seed_dt=0.1):
    return W.desc(w, h, [dict(radius=float(radius), beta=[1.0, 0.5], mu=0.2, sigma=0.03)], seed_dt, "fft_fp32",
                  init=W.INIT_PATCH, init_size=min(w, h) // 4)


@pytest.mark.parametrize("kernels", ["fused", "block", "generic", "tune"])
def test_batch_fft_group_shapes(O, monkeypatch, kernels):
    """FFT groups at the edges of the batch kernels' ranges: heights 64 and
    2048 (the fused column pass with four and two columns per workgroup),
    256, 512, 1024 and 2048 (the block column pass, Q = 1, 2, 4, 8), height
    4096 (the generic column pass in diagonal mode), width 64 (the generic
    row pass) and widths 128, 256, 512, 1024 and 2048 (the fused row pass),
    two worlds per group.  Run four ways: the fused kernels wherever they
    apply, the block column kernel wherever it applies, the generic kernels
    everywhere, and the tuner's choice per size.  Every world vs a
    standalone Sim (within 1e-5) and vs the FP64 step (within 1e-3) from the
    same input."""
    monkeypatch.setenv("AUTO_HIP_BATCH_FFT_KERNELS", kernels)
    shapes = [(64, 64, 8), (2048, 2048, 20), (64, 4096, 12), (2048, 64, 12), (128, 2048, 16),
              (512, 512, 18), (1024, 1024, 24), (256, 256, 14), (64, 512, 10)]
    worlds = [(_world_wh(w, h, r), 10 * i + j) for i, (w, h, r) in enumerate(shapes) for j in range(2)]
    b = Batch(worlds)
    info = b.info()
    assert info.fft_worlds == len(worlds) and info.fft_groups == len(shapes)
    if kernels == "tune":
        # Every group timed at least once (the generic pair always applies),
        # and a group where every batch kernel applies times all six pairs
        # (column: 0 fused, 1 generic, 2 block; row: 0 fused, 1 generic).
        seen = {}
        for k in range(info.n_group_tune):
            t = info.group_tune[k]
            assert t.ms_per_step > 0
            seen.setdefault((t.width, t.height), set()).add((t.column, t.row))
        assert set(seen) == {(w, h) for w, h, _ in shapes}
        assert all((1, 1) in v for v in seen.values())
        assert seen[(512, 512)] == {(c, r) for c in (0, 1, 2) for r in (0, 1)}
    else:
        assert info.n_group_tune == 0
    b.step(2)
    before = [b.read_world(i) for i in range(len(worlds))]
    b.step(1)
    worst64 = worst_sim = 0.0
    for i, (d, seed) in enumerate(worlds):
        got = b.read_world(i)
        worst64 = max(worst64, float(np.abs(got - O.step64(d, before[i][None], "fft")[0]).max()))
        sim = Sim(d, seed)
        sim.write_field("state", before[i][None])
        sim.step(1)
        worst_sim = max(worst_sim, float(np.abs(got - sim.read_field("state")[0]).max()))
        assert got.std() > 0, i  # the patch is alive
    print(f"group shapes: worst max-abs vs FP64 {worst64:.3e}, vs standalone Sims {worst_sim:.3e}")
    assert worst64 <= 1e-3 and worst_sim < 1e-5


def test_batch_group_split_and_order_are_bitwise_neutral(monkeypatch):
    """FFT worlds of one size split into several groups (a 1 MiB working-set
    cap), stepped group-major (the default) and step-major, through graph
    replays and eager steps: every world's state is identical bit for bit
    to the unsplit step-major batch.  The fused kernels are forced, so the
    tuner's choice cannot differ between the batches."""
    monkeypatch.setenv("AUTO_HIP_BATCH_FFT_KERNELS", "fused")
    worlds = W.uneven_worlds(seed=11, n=14, path="fft_fp32", sides=(128, 256))
    worlds += W.uneven_worlds(seed=12, n=3, path="direct_bf16", sides=(128,))
    runs = []
    for cap, order in (("0", "step"), ("1", "group"), ("1", "step"), ("0", "group")):
        monkeypatch.setenv("AUTO_HIP_BATCH_GROUP_MB", cap)
        monkeypatch.setenv("AUTO_HIP_BATCH_ORDER", order)
        b = Batch(worlds)
        groups = b.info().fft_groups
        b.step(103)  # one 100-step graph replay, then three eager steps
        runs.append(([b.read_world(i) for i in range(len(worlds))], groups))
    sizes = {(d.width, d.height) for d, _ in worlds[:14]}
    assert runs[0][1] == len(sizes) and runs[1][1] > len(sizes) + 2
    for states, _ in runs[1:]:
        for i in range(len(worlds)):
            assert np.array_equal(states[i], runs[0][0][i]), i
    assert any(runs[0][0][i].std() > 0 for i in range(len(worlds)))


def _bench_batch_vs_singles(O, worlds, b):
    """Every world of a stepped batch against a standalone FFT Sim and the
    FP64 oracle step from the same input state: (worst vs Sim, worst vs FP64)."""
    before = [b.read_world(i) for i in range(len(worlds))]
    b.step(1)
    worst_sim = worst64 = 0.0
    where = None
    for i, (d, seed) in enumerate(worlds):
        got = b.read_world(i)
        plain = LeniaDesc.from_buffer_copy(d)
        plain.path = PATH_FFT_FP32
        sim = Sim(plain, seed)
        sim.write_field("state", before[i][None])
        sim.step(1)
        e = float(np.abs(got - sim.read_field("state")[0]).max())
        if e > worst_sim:
            worst_sim, where = e, (i, d.width)
        sim.close()
        worst64 = max(worst64, float(np.abs(got - O.step64(d, before[i][None], "fft")[0]).max()))
    return worst_sim, worst64, where


def test_benchmark_batch_384_auto(O):
    """The 384-world batch exactly as the benchmark builds it (seed 1, sides
    {128, 256, 512, 1024}, path AUTO): after 2 steps, every world's next step
    matches a standalone FFT Sim from the same input (within 1e-5) and the
    FP64 oracle step (within 1e-3).  All four size groups are present, so
    the fused column pass runs at heights 128..1024 and the diagonal row
    pass at widths 128..1024."""
    worlds = W.uneven_worlds(seed=1, n=384, path="auto")
    b = Batch(worlds)
    info = b.info()
    sides = {d.width for d, _ in worlds}
    assert sides == {128, 256, 512, 1024}
    print(f"benchmark batch: {info.fft_worlds} FFT worlds in {info.fft_groups} groups, {info.ntiles} direct tiles")
    b.step(2)
    worst_sim, worst64, where = _bench_batch_vs_singles(O, worlds, b)
    print(f"384 worlds: worst max-abs vs standalone FFT Sims {worst_sim:.3e} (world, side {where}), "
          f"vs FP64 {worst64:.3e}")
    assert worst_sim < 1e-5 and worst64 <= 1e-3


@pytest.mark.parametrize("kernels", ["fused", "block"])
def test_benchmark_batch_384_graph_replays_equal_eager(monkeypatch, kernels):
    """The benchmark batch stepped by graph replays (100 steps each) equals
    the same batch stepped by single eager launches, bit for bit, after the
    first replay and after a later one with eager work in between.  The
    column kernels are forced (two batches tuned apart may time their
    candidates differently and pick different kernels)."""
    monkeypatch.setenv("AUTO_HIP_BATCH_FFT_KERNELS", kernels)
    worlds = W.uneven_worlds(seed=1, n=384, path="auto")
    a, b = Batch(worlds), Batch(worlds)
    for rnd in range(2):
        a.step(100)
        for _ in range(100):
            b.step(1)
        a.wait()
        b.wait()
        for i in range(len(worlds)):
            assert np.array_equal(a.read_world(i), b.read_world(i)), (rnd, i)
