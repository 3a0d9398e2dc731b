#!/usr/bin/env python3
"""Lenia world-step benchmark on one MI355X.

Workload: the 4096x4096 three-ring world (R=26, beta=[1,2/3,1/3], mu=0.26,
sigma=0.036, dt=0.1, 64 random 48x48 blobs), FP32 state.  One step is one
Lenia generation of the whole torus; cell-updates/s = width*height*steps/s as
the reference computes it (crates/auto-cli/src/bench.rs:161).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--path fft_fp32]

Prints one JSON line.  `value` is the FFT path's cell-updates/s with the state
resident in HBM: the median over repeated blocks of exactly K steps, each
bracketed by stream synchronisation (the reference's calibrate-then-median
method, crates/auto-cli/src/bench.rs:126-158), with the 10-90% spread beside
it.  Per-kernel device times come from HIP events on the engine's own stream
in a separate window of at least 256 samples per kernel; `roofline` prices the
dominant kernel's algorithmic bytes against the 8 TB/s HBM peak.  The CPU
baseline times the oracle's FP32 restatement of the reference step on host
cores (one thread and all threads, 256^2 and 4096^2, calibrate-then-median),
and the error legs check GPU steps against the oracle's FP64 step from the
same input along the runs.  The 384-world batch is timed beside the same
worlds stepped one by one as standalone sims.

Opt-in comparisons with the library route on the same GPU (each spends
minutes in MIOpen's algorithm search on a fresh machine): --torch-conv (the
direct path's 53x53 convolution as torch conv2d) and --torch-surrogate (the
surrogate's network trained in PyTorch).  --surrogate-only
[--surrogate-radius R] runs the surrogate section alone.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# Timing method (crates/auto-cli/src/bench.rs:126-158 grows the batch to
# >= 20 ms and takes the median over >= 3 runs): warm up for WARMUP_S, then
# time blocks of exactly --steps steps, each bracketed by stream
# synchronisation, for at least TIME_S and MIN_BLOCKS blocks; `value` is the
# median block's rate, with the 10-90% spread beside it.  Kernel durations
# come from a separate window after the timed blocks, with HIP events around
# every EVENT_EVERY-th launch of each kernel, EVENT_SAMPLES samples each.
WARMUP_S = 0.5
TIME_S = 1.0
MIN_BLOCKS = 3
REF_BATCH_S = 0.020
EVENT_EVERY = 4
EVENT_SAMPLES = 256

# HBM bytes per launch that rocprofv3 measured for the dominant kernels of
# the 4096^2 three-ring world (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the
# factors a copy kernel of known size reads; summary and commands in
# profiles/fft_fp32_rings3_4096.md and profiles/direct_bf16_rings3_4096.md).
# A bench run cannot collect counters itself, so `roofline.traffic` reports
# this recorded figure for that world and null for any other size.
RECORDED_TRAFFIC = {
    ("row_c2r_update_r2c", 4096, 4096): 275952558.0,
    ("col_conv", 4096, 4096): 151500206.0,
    ("direct_bf16_update", 4096, 4096): 168327361.0,
}
TRAFFIC_SOURCE = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, recorded in profiles/"

# FFT-path kernels priced by algorithmic HBM bytes (single channel, one
# kernel; the half spectrum is W/2 x H complex64 = 4 B/cell).
FFT_KERNELS = ("row_r2c", "col_conv", "row_c2r_update", "row_c2r_update_r2c")


def khat_bytes(w: int, h: int) -> int:
    """K-hat bytes one col_conv launch reads: h/2+1 floats for each of the
    w/2 columns (the spectrum is real and even in ky), plus the packed
    column's h/2+1 float2."""
    return (w // 2) * (h // 2 + 1) * 4 + (h // 2 + 1) * 8


def kernel_bytes(name: str, w: int, h: int) -> float:
    """row_r2c: read state + write spectrum; col_conv: read + write spectrum
    + K-hat; row_c2r_update: read spectrum + read state + write state;
    row_c2r_update_r2c: that plus writing the next step's spectrum."""
    cells = w * h
    if name == "row_r2c":
        return 8.0 * cells
    if name == "col_conv":
        return 8.0 * cells + khat_bytes(w, h)
    if name == "row_c2r_update":
        return 12.0 * cells
    if name == "row_c2r_update_r2c":
        return 16.0 * cells
    raise KeyError(name)


def workload_name(d) -> str:
    k = d.kernels[0]
    return (f"rings3_{d.width}: {d.width}x{d.height} torus, R={k.radius:g}, "
            f"beta={[round(b, 4) for b in list(k.beta)[:k.rings]]}, mu={k.mu:g}, sigma={k.sigma:g}, "
            f"dt={d.dt:g}, {d.init_count} blobs of {d.init_size}^2")


def dense_box_flops_per_cell(d) -> float:
    """Algorithmic flops of a direct convolution per cell: 2 per tap of the
    dense (2*reach+1)^2 box, reach = ceil(R) (2*53^2 = 5,618 at R=26)."""
    r = math.ceil(d.kernels[0].radius)
    return 2.0 * (2 * r + 1) ** 2


def useful_flops_per_cell(d) -> float:
    """2 flops per nonzero tap of the kernel (0 < r < R)."""
    R = d.kernels[0].radius
    n = int(R) + 1
    return 2.0 * sum(1 for y in range(-n, n + 1) for x in range(-n, n + 1) if 0 < (x * x + y * y) ** 0.5 < R)


def log(msg: str) -> None:
    """Progress on stderr; stdout carries only the JSON line."""
    print(msg, file=sys.stderr, flush=True)


def quantile(xs, q: float) -> float:
    ys = sorted(xs)
    if not ys:
        return float("nan")
    pos = q * (len(ys) - 1)
    lo = int(math.floor(pos))
    hi = min(lo + 1, len(ys) - 1)
    return ys[lo] + (ys[hi] - ys[lo]) * (pos - lo)


def warm_up(step, wait, steps: int, warmup: int) -> None:
    """At least `warmup` steps and at least WARMUP_S seconds of stepping."""
    step(max(warmup, 1))
    wait()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < WARMUP_S:
        step(max(steps, 16))
        wait()


def timed_blocks(step, wait, steps: int, cells: int, min_s: float = TIME_S, min_blocks: int = MIN_BLOCKS):
    """Blocks of exactly `steps` steps, each bracketed by synchronisation,
    until min_s seconds and min_blocks blocks have run.  The median block
    gives the rate; p10/p90 of the per-block rates give the spread."""
    secs = []
    t_start = time.perf_counter()
    while len(secs) < min_blocks or time.perf_counter() - t_start < min_s:
        t0 = time.perf_counter()
        step(steps)
        wait()
        secs.append(time.perf_counter() - t0)
    rates = [cells * steps / s for s in secs]
    med = quantile(secs, 0.5)
    return {"value": cells * steps / med, "ms_per_step": 1e3 * med / steps, "blocks": len(secs),
            "block_steps": steps, "p10": quantile(rates, 0.1), "p90": quantile(rates, 0.9)}


def reference_method(step, wait, cells: int):
    """crates/auto-cli/src/bench.rs:126-158: double the batch until one run
    takes >= 20 ms, then the median over >= 3 runs (and >= 1 s)."""
    n = 1
    while True:
        t0 = time.perf_counter()
        step(n)
        wait()
        if time.perf_counter() - t0 >= REF_BATCH_S:
            break
        n *= 2
    r = timed_blocks(step, wait, n, cells)
    return {"batch_steps": n, "value": r["value"], "ms_per_step": r["ms_per_step"], "runs": r["blocks"],
            "p10": r["p10"], "p90": r["p90"]}


def event_window(sim, steps_hint: int = 1):
    """Per-kernel device time from HIP events on the sim's stream around
    every EVENT_EVERY-th launch of each kernel, over enough steps that every
    kernel launched once per step has EVENT_SAMPLES samples."""
    window = EVENT_EVERY * EVENT_SAMPLES
    sim.profile(True, every=EVENT_EVERY)
    sim.step(window)
    sim.wait()
    prof = sim.profile_read()
    sim.profile(False)
    out = {}
    for n, (launches, ms) in prof.items():
        out[n] = {"timed_launches": launches, "avg_us": 1e3 * ms / launches,
                  "launches_per_step": launches * EVENT_EVERY / window}
    return out


def time_path(path: str, side: int, steps: int, warmup: int, seed: int):
    from auto_amd import Sim, workloads as W
    d = W.rings3_4096(path=path, side=side)
    sim = Sim(d, seed)
    cells = d.width * d.height
    warm_up(sim.step, sim.wait, steps, warmup)
    r = timed_blocks(sim.step, sim.wait, steps, cells)
    r["reference_method"] = reference_method(sim.step, sim.wait, cells)
    r["kernels"] = event_window(sim)
    return d, sim, r


def batch_profile_run(seed: int, steps: int, n: int = 384) -> dict:
    """The benchmark batch (AUTO) for a profiler: 100 graph-replayed warm-up
    steps, then one timed block of `steps` steps (rounded up to whole
    100-step graphs)."""
    from auto_amd import Batch, workloads as W
    b = Batch(W.uneven_worlds(seed=seed, n=n, path="auto"))
    info = b.info()
    b.step(100)
    b.wait()
    steps = 100 * max(1, (steps + 99) // 100)
    t0 = time.perf_counter()
    b.step(steps)
    b.wait()
    t = time.perf_counter() - t0
    return {"worlds": n, "fft_groups": info.fft_groups, "steps": steps, "ms_per_step": 1e3 * t / steps,
            "value": info.total_cells * steps / t, "unit": "cell-updates/s",
            "fft_group_tune": [{"world": [g.width, g.height], "column": ("fused", "generic", "block")[g.column],
                                "row": "generic" if g.row else "fused", "ms_per_step": g.ms_per_step}
                               for g in info.group_tune[:info.n_group_tune]]}


def time_batch(seed: int, n: int = 384, graph_steps: int = 100, path: str = "auto", check: bool = True,
               measure: bool = True):
    """The 384 uneven worlds: aggregate cell-updates/s over graph-replayed
    steps (median of 100-step blocks), autotuner choices, the launch-overhead
    share, and each world's error against the plain single-world FFT step
    from the same input.  path "auto" lets the batch put each world on the
    FFT or the direct path by its cost model; "direct_bf16" keeps every world
    on the direct path."""
    import numpy as np
    from auto_amd import Batch, Sim, workloads as W
    from auto_amd._abi import LeniaDesc, PATH_FFT_FP32
    worlds = W.uneven_worlds(seed=seed, n=n, path=path)
    b = Batch(worlds)
    info = b.info()
    cells = info.total_cells
    b.step(graph_steps)  # warm-up: captures the graph
    b.wait()
    r = timed_blocks(b.step, b.wait, graph_steps, cells)
    m = b.measure(graph_steps) if measure else None
    out = {"worlds": n, "path": path, "fft_worlds": info.fft_worlds, "fft_groups": info.fft_groups,
           "cells_per_step": cells, "value": r["value"], "unit": "cell-updates/s", "ms_per_step": r["ms_per_step"],
           "p10": r["p10"], "p90": r["p90"], "blocks": r["blocks"],
           "tile": [info.tile_cols, info.tile_rows], "grid": info.grid, "tiles": info.ntiles,
           "autotune": [{"tile": [info.tune[i].tile_cols, info.tune[i].tile_rows], "grid": info.tune[i].grid,
                         "ms_per_step": info.tune[i].ms_per_step} for i in range(info.n_tune)],
           "fft_group_tune": [{"world": [t.width, t.height], "column": ("fused", "generic", "block")[t.column],
                               "row": "generic" if t.row else "fused", "ms_per_step": t.ms_per_step}
                              for t in info.group_tune[:info.n_group_tune]],
           "graph_steps": graph_steps,
           # 1 - (summed device time of the launches, from per-launch events
           # on eager steps) / (wall time of graph replays of as many steps).
           # Per-launch events lengthen each timed launch slightly, so the
           # share can come out at or below 0: then it is reported as 0.
           "launch_overhead_share": max(0.0, 1.0 - m["kernel_ms"] / m["graph_wall_ms"]) if m else None,
           "eager_over_graph_wall": m["eager_wall_ms"] / m["graph_wall_ms"] if m else None,
           "measure": m}
    if check:
        before = [b.read_world(i) for i in range(n)]
        b.step(1)
        worst, worst_rel, where = 0.0, 0.0, None
        for i, (d, s) in enumerate(worlds):
            got = b.read_world(i)
            plain = LeniaDesc.from_buffer_copy(d)
            plain.path = PATH_FFT_FP32
            sim = Sim(plain, s)
            sim.write_field("state", before[i][None])
            sim.step(1)
            ref = sim.read_field("state")[0]
            sim.close()
            e = float(np.abs(got - ref).max())
            nrm = float(np.linalg.norm(ref))
            if e >= worst:
                worst, where = e, {"world": i, "side": d.width}
            if nrm > 0:
                worst_rel = max(worst_rel, float(np.linalg.norm(got - ref)) / nrm)
        out["per_world_vs_plain_step"] = {"worst_max_abs": worst, "worst_rel_l2": worst_rel, "at": where,
                                          "protocol": "one batch step vs one plain single-world FFT step "
                                                      "from the same input, every world"}
    return out


def time_plain_worlds(seed: int, n: int = 384, steps: int = 20):
    """The batch's worlds without the batch: every world a standalone FFT Sim
    (the plain single-world step, the way a caller of Sim::step advances one
    world at a time, crates/auto-core/src/sim.rs:297-309), each block queueing
    `steps` steps on every world's own stream before waiting for all of them,
    so worlds on different streams may overlap on the GPU.  Median of blocks
    as timed_blocks does."""
    from auto_amd import Sim, workloads as W
    from auto_amd._abi import LeniaDesc, PATH_FFT_FP32
    sims, cells = [], 0
    t0 = time.perf_counter()
    for d, s in W.uneven_worlds(seed=seed, n=n):
        plain = LeniaDesc.from_buffer_copy(d)
        plain.path = PATH_FFT_FP32
        sims.append(Sim(plain, s))
        cells += d.width * d.height
    setup_s = time.perf_counter() - t0

    def step(k):
        for sim in sims:
            sim.step(k)

    def wait():
        for sim in sims:
            sim.wait()

    step(2)
    wait()
    r = timed_blocks(step, wait, steps, cells)
    for sim in sims:
        sim.close()
    return {"worlds": n, "cells_per_step": cells, "value": r["value"], "unit": "cell-updates/s",
            "ms_per_step": r["ms_per_step"], "p10": r["p10"], "p90": r["p90"], "blocks": r["blocks"],
            "block_steps": steps, "setup_s": setup_s,
            "method": f"{n} standalone FFT Sims, {steps} steps queued on each world's stream per block, "
                      "then a wait on every stream; median block"}


def run_4096(seed: int, steps: int = 1000, every: int = 100):
    """Config-2 error protocol along the run: the 4096^2 three-ring world on
    both paths for `steps` steps; every `every` steps, one step from the
    current state against the oracle's FP64 step from the same input
    (max-abs, rel-L2).  Mass drift: the direct path's mass against the FFT
    path's at the end (the FFT path tracks the FP64 trajectory's mass to
    ~2e-5 over 1000 steps at 1024^2, tests/test_gpu_drift.py)."""
    import numpy as np
    from auto_amd import Sim, workloads as W
    from oracle import oracle as O
    out = {"steps": steps, "every": every}
    masses = {}
    for path in ("fft_fp32", "direct_bf16"):
        d = W.rings3_4096(path=path)
        sim = Sim(d, seed)
        m0 = sim.stats().mass[0]
        samples = []
        t = 0
        while t < steps:
            s = sim.read_field("state")
            sim.step(1)
            got = sim.read_field("state")
            want = O.step64(d, s, "fft")
            diff = got.astype(np.float64) - want
            nrm = float(np.linalg.norm(want))
            samples.append({"t": t, "max_abs": float(np.abs(diff).max()),
                            "rel_l2": float(np.linalg.norm(diff)) / nrm if nrm > 0 else None})
            sim.step(every - 1)
            t += every
        masses[path] = (m0, sim.stats().mass[0])
        out[path] = {"samples": samples, "worst_max_abs": max(x["max_abs"] for x in samples),
                     "worst_rel_l2": max((x["rel_l2"] or 0.0) for x in samples)}
        log(f"run_4096 {path}: worst one-step max-abs vs FP64 over {len(samples)} samples "
            f"{out[path]['worst_max_abs']:.2e}")
    f, dm = masses["fft_fp32"][1], masses["direct_bf16"][1]
    out["mass0"] = masses["fft_fp32"][0]
    out["mass_fft"], out["mass_direct"] = f, dm
    out["mass_drift_direct_vs_fft"] = abs(dm - f) / f
    return out


def time_rule(rule: str, side: int, steps: int, seed: int):
    """SmoothLife (ra = 21) or MNCA (rings [12,18), [3,8)) on the FFT path:
    two kernels, so the column pass writes two product spectra and the row
    pass inverts both before the transition."""
    from auto_amd import Sim, workloads as W
    d = W.smoothlife(side, side) if rule == "smoothlife" else W.mnca(side, side)
    sim = Sim(d, seed)
    warm_up(sim.step, sim.wait, steps, 10)
    r = timed_blocks(sim.step, sim.wait, steps, d.width * d.height)
    return {"value": r["value"], "unit": "cell-updates/s", "ms_per_step": r["ms_per_step"],
            "p10": r["p10"], "p90": r["p90"],
            "workload": f"{rule} {d.width}x{d.height}, reference parameters, blocks of {steps} steps",
            "kernels": event_window(sim)}


def time_multichannel(path: str, steps: int, seed: int):
    """1024^2 x 3 channels, 9 kernels (k = 3*src + dst) drawn from SplitMix64."""
    from auto_amd import Sim, workloads as W
    d = W.multichannel_1024(seed=seed, path=path)
    sim = Sim(d, seed)
    warm_up(sim.step, sim.wait, steps, 10)
    cells = d.width * d.height
    r = timed_blocks(sim.step, sim.wait, steps, cells)
    return d, sim, {"value": r["value"], "unit": "cell-updates/s (3 channels per cell)",
                    "channel_cell_updates_per_s": 3 * r["value"], "ms_per_step": r["ms_per_step"],
                    "p10": r["p10"], "p90": r["p90"], "kernels": event_window(sim)}


def multichannel_errors(seed: int, steps: int = 500, every: int = 50):
    """Config-3 error protocol: per channel, one step from the current state
    against the oracle's FP64 step from the same input, every `every` steps
    over `steps` steps, on both paths; with each sample's live cells."""
    import numpy as np
    from auto_amd import Sim, workloads as W
    from oracle import oracle as O
    out = {"steps": steps, "every": every}
    for path in ("fft_fp32", "direct_bf16"):
        d = W.multichannel_1024(seed=seed, path=path)
        sim = Sim(d, seed)
        samples = []
        t = 0
        while t < steps:
            s = sim.read_field("state")
            sim.step(1)
            got = sim.read_field("state")
            want = O.step64(d, s, "fft")
            per = []
            for c in range(d.channels):
                diff = got[c].astype(np.float64) - want[c]
                nrm = float(np.linalg.norm(want[c]))
                per.append({"max_abs": float(np.abs(diff).max()),
                            "rel_l2": float(np.linalg.norm(diff)) / nrm if nrm > 0 else None,
                            "live_cells": int(np.count_nonzero(want[c]))})
            samples.append({"t": t, "channels": per})
            sim.step(every - 1)
            t += every
        out[path] = {"samples": samples,
                     "worst_max_abs_per_channel": [max(x["channels"][c]["max_abs"] for x in samples)
                                                   for c in range(d.channels)]}
        log(f"multichannel {path}: worst per-channel one-step max-abs vs FP64 "
            f"{[f'{v:.2e}' for v in out[path]['worst_max_abs_per_channel']]}")
    return out


def time_torch_fft(side: int, steps: int, seed: int):
    """The same FP32 step as plain PyTorch eager code on the GPU, the way a
    user without this engine would write it: torch.fft.rfft2 / irfft2 (rocFFT)
    and elementwise growth.  The kernel comes from the engine itself (the
    potential of a unit impulse at the origin is the kernel), so both sides
    convolve with the same taps.  PyTorch runs in a child process: its own
    HIP runtime sees no device in a process where this engine's runtime
    already holds it."""
    import subprocess
    import tempfile
    import numpy as np
    from auto_amd import Sim, workloads as W
    d = W.rings3_4096(side=side)
    k = d.kernels[0]
    probe = Sim(d, seed)
    impulse = np.zeros((1, d.height, d.width), np.float32)
    impulse[0, 0, 0] = 1.0
    probe.write_field("state", impulse)
    kern = probe.read_field("potential0").reshape(d.height, d.width)
    del probe
    ours = Sim(d, seed)
    ours.step(20)
    s0 = ours.read_field("state").reshape(d.height, d.width)
    ours.step(1)
    one = ours.read_field("state").reshape(d.height, d.width)
    del ours
    with tempfile.TemporaryDirectory() as tmp:
        np.savez(os.path.join(tmp, "in.npz"), kern=kern, s0=s0, one=one,
                 params=np.array([k.mu, k.sigma, d.dt, k.weight], np.float64), steps=np.array(steps))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--torch-leg", tmp],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"torch leg failed ({r.returncode}): {r.stderr[-2000:]}")
        return json.loads(r.stdout.strip().splitlines()[-1])


def torch_leg(tmp: str) -> int:
    """Child process of time_torch_fft: PyTorch only, this engine not loaded."""
    import numpy as np
    import torch
    z = np.load(os.path.join(tmp, "in.npz"))
    mu, sigma, dt, weight = (float(v) for v in z["params"])
    steps = int(z["steps"])
    dev = torch.device("cuda:0")
    khat = torch.fft.rfft2(torch.from_numpy(z["kern"]).to(dev))

    def step(s):
        u = torch.fft.irfft2(torch.fft.rfft2(s) * khat, s=s.shape)
        g = 2.0 * torch.exp(-0.5 * ((u - mu) / sigma) ** 2) - 1.0
        return torch.clamp(s + (dt * weight) * g, 0.0, 1.0)

    s = torch.from_numpy(z["s0"]).to(dev)
    err = float((step(s) - torch.from_numpy(z["one"]).to(dev)).abs().max())
    for _ in range(10):
        s = step(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s = step(s)
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    cells = s.numel()
    print(json.dumps({"value": cells * steps / dt_s, "unit": "cell-updates/s", "ms_per_step": 1e3 * dt_s / steps,
                      "steps": steps, "torch": torch.__version__, "max_abs_vs_engine_one_step": err}), flush=True)
    return 0


def time_torch_conv(side: int = 4096, radius: int = 26, steps: int = 20):
    """The direct path's convolution through the library: torch conv2d
    (MIOpen) of a side^2 torus with one (2 radius + 1)^2 kernel, circular
    padding, bf16 and fp32, ms per convolution (the convolution alone, no
    growth or update) after 3 warm-up calls (MIOpen's algorithm search takes
    minutes on a fresh machine).  Child process, as time_torch_fft."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--torch-conv-leg", f"{side},{radius},{steps}"],
                       stdout=subprocess.PIPE, text=True, timeout=1100)
    if r.returncode != 0:
        raise RuntimeError(f"torch conv leg failed ({r.returncode})")
    return json.loads(r.stdout.strip().splitlines()[-1])


def torch_conv_leg(spec: str) -> int:
    """Child process of time_torch_conv: PyTorch only."""
    import torch
    import torch.nn.functional as F
    side, radius, steps = (int(v) for v in spec.split(","))
    dev = torch.device("cuda:0")
    out = {"side": side, "kernel": 2 * radius + 1, "torch": torch.__version__}
    for name, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
        x = torch.rand(1, 1, side, side, device=dev, dtype=dt)
        k = torch.rand(1, 1, 2 * radius + 1, 2 * radius + 1, device=dev, dtype=dt)

        def conv():
            return F.conv2d(F.pad(x, (radius, radius, radius, radius), mode="circular"), k)

        t0 = time.perf_counter()
        for i in range(3):
            conv()
            torch.cuda.synchronize()
            log(f"torch conv2d {name} warm-up call {i + 1}/3 done at {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        for _ in range(steps):
            conv()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        out[name] = {"ms_per_conv": ms, "cell_updates_per_s": side * side / (ms * 1e-3)}
    print(json.dumps(out), flush=True)
    return 0


def time_torch_surrogate(batch: int = 64, side: int = 256, steps: int = 30):
    """The surrogate's network written in PyTorch (torch.nn.Conv2d on MIOpen,
    bf16 weights and activations, circular padding, exact GELU, 2 x 2 average
    pool, nearest upsampling, the skip and the clipped residual), one Adam
    step on a batch of random states: ms per step after 5 warm-up steps (the
    first ones run MIOpen's algorithm search, which takes minutes on a fresh
    machine).  Child process, as time_torch_fft."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--torch-surrogate-leg",
                        f"{batch},{side},{steps}"], stdout=subprocess.PIPE, text=True, timeout=1100)
    if r.returncode != 0:
        raise RuntimeError(f"torch surrogate leg failed ({r.returncode})")
    return json.loads(r.stdout.strip().splitlines()[-1])


def torch_surrogate_leg(spec: str) -> int:
    """Child process of time_torch_surrogate: PyTorch only."""
    import torch
    import torch.nn.functional as F
    batch, side, steps = (int(v) for v in spec.split(","))
    dev = torch.device("cuda:0")
    torch.manual_seed(0)

    class Net(torch.nn.Module):
        def __init__(self):
            super().__init__()
            self.c1 = torch.nn.Conv2d(1, 32, 5)
            self.c2 = torch.nn.Conv2d(32, 64, 5)
            self.c3 = torch.nn.Conv2d(64, 32, 5)
            self.c4 = torch.nn.Conv2d(32, 1, 5)

        def forward(self, x):
            def pad(t):
                return F.pad(t, (2, 2, 2, 2), mode="circular")
            a1 = F.gelu(self.c1(pad(x)))
            a2 = F.gelu(self.c2(pad(F.avg_pool2d(a1, 2))))
            s = F.gelu(self.c3(pad(F.interpolate(a2, scale_factor=2, mode="nearest")))) + a1
            return torch.clamp(x + self.c4(pad(s)), 0.0, 1.0)

    net = Net().to(dev).to(torch.bfloat16)
    params = sum(p.numel() for p in net.parameters())
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    x = torch.rand(batch, 1, side, side, device=dev, dtype=torch.bfloat16)
    t = torch.rand_like(x)

    def step():
        opt.zero_grad(set_to_none=True)
        loss = F.mse_loss(net(x).float(), t.float())
        loss.backward()
        opt.step()

    t0 = time.perf_counter()
    for i in range(5):
        step()
        torch.cuda.synchronize()
        log(f"torch surrogate warm-up step {i + 1}/5 done at {time.perf_counter() - t0:.1f} s")
    warm_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    print(json.dumps({"train_ms_per_step": 1e3 * dt_s / steps, "batch": batch, "crop": side, "steps": steps,
                      "params": params, "warmup_s": warm_s, "torch": torch.__version__,
                      "dtype": "bf16 weights and activations"}), flush=True)
    return 0


def time_surrogate(seed: int, train_steps: int, pairs: int = 20000, crop: int = 256, steps: int = 1000,
                   radius: float = 0.0):
    """Learned 4-step surrogate: pairs from the 4096^2 three-ring trajectory,
    Adam training, then `steps` judged steps at 4096^2 beside plain stepping.
    With `radius` set, the world is instead a 1024^2 one-ring world of that
    radius (the three-ring world's mu, sigma and dt, seeded with uniform noise
    everywhere), so that 4 steps depend on a neighbourhood the network's
    receptive field (about 11 cells) covers when radius <= ~3."""
    import numpy as np
    from auto_amd import Sim, Surrogate, workloads as W
    if radius > 0:
        d = W.desc(1024, 1024, [dict(radius=float(radius), beta=[1.0], mu=0.26, sigma=0.036)], 0.1,
                   init=W.INIT_PATCH, init_size=1024)
    else:
        d = W.rings3_4096()
    gen = Sim(d, seed + 100)
    gen.step(50)
    net = Surrogate(seed=seed)
    t0 = time.perf_counter()
    net.collect(gen, pairs=pairs, crop=crop, horizon=4, seed=seed)
    t1 = time.perf_counter()
    log(f"surrogate: collected {pairs} pairs of {crop}^2 in {t1 - t0:.1f} s")
    held = min(512, pairs // 10)
    # Chunks of 100 steps (Adam state carries over between calls) so the
    # run reports progress; each chunk samples batches from its own seed.
    losses = []
    for c, done in enumerate(range(0, train_steps, 100)):
        n = min(100, train_steps - done)
        losses += list(net.train(steps=n, batch=64, n_train=pairs - held, lr=1e-3, seed=seed * 1000 + c))
        log(f"surrogate: {done + n}/{train_steps} training steps, last-{n} mean loss "
            f"{float(np.mean(losses[-n:])):.3e}, {time.perf_counter() - t1:.1f} s")
    t2 = time.perf_counter()
    held_mse = net.eval(pairs - held, held)
    xh, yh = net.dataset(pairs - held, held)
    a, b = Sim(d, seed), Sim(d, seed)
    a.step(50)
    b.step(50)
    log(f"surrogate: held-out MSE {held_mse:.3e}; advancing {steps} steps at {d.width}^2")
    rep = net.advance(a, steps, judge_every=16, tol=1e-2)
    t3 = time.perf_counter()
    b.step(steps)
    b.wait()
    t4 = time.perf_counter()
    sa, sb = a.read_field("state"), b.read_field("state")
    cells = d.width * d.height
    return {"world": f"{d.width}x{d.height}, R={d.kernels[0].radius:g}, {d.kernels[0].rings} ring(s)",
            "pairs": pairs, "crop": crop, "collect_s": t1 - t0, "train_steps": train_steps, "batch": 64,
            "train_s": t2 - t1, "train_ms_per_step": 1e3 * (t2 - t1) / max(train_steps, 1),
            "loss_first10": float(np.mean(losses[:10])), "loss_last10": float(np.mean(losses[-10:])),
            "held_out_mse": held_mse, "held_out_identity_mse": float(np.mean((xh - yh) ** 2)),
            "advance": rep, "value": cells * steps / (rep["ms"] * 1e-3), "unit": "cell-updates/s",
            "plain_value": cells * steps / (t4 - t3),
            "final_rel_l2_vs_plain": float(np.linalg.norm(sa - sb) / np.linalg.norm(sb)),
            "mass_drift_vs_plain": float(abs(sa.sum(dtype=np.float64) - sb.sum(dtype=np.float64))
                                         / sb.sum(dtype=np.float64))}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        return ""


def cpu_rate(O, d, s, spectra, threads: int, budget_s: float):
    """The reference's method on the oracle's FP32 step: double the batch
    until one run takes >= 20 ms, then runs of that batch until budget_s and
    >= 3 runs; median rate and 10-90% spread."""
    O.set_threads(threads)
    cells = d.width * d.height
    n = 1
    while True:
        t0 = time.perf_counter()
        x = s
        for _ in range(n):
            x = O.step_ref32(d, x, spectra)
        if time.perf_counter() - t0 >= REF_BATCH_S:
            break
        n *= 2
    secs = []
    t_start = time.perf_counter()
    while len(secs) < 3 or time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        x = s
        for _ in range(n):
            x = O.step_ref32(d, x, spectra)
        secs.append(time.perf_counter() - t0)
    rates = [cells * n / t for t in secs]
    return {"value": cells * n / quantile(secs, 0.5), "threads": threads, "batch_steps": n, "runs": len(secs),
            "seconds": sum(secs), "p10": quantile(rates, 0.1), "p90": quantile(rates, 0.9)}


def cpu_baseline(d4096, s4096, budget_s: float):
    """The oracle's C restatement of the reference FP32 Stockham step (kind
    "port": the reference has no CPU Lenia) at 256^2 (config 1) and 4096^2
    (config 2), on one thread and on all host threads.  `value` is 4096^2 on
    all threads."""
    from auto_amd import workloads as W
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    rows = {}
    d256 = W.single_ring_256()
    s256 = O.init_state(d256, 1)
    for name, d, s, budget in (("256", d256, s256, 2.0), ("4096", d4096, s4096, budget_s)):
        spectra = O.kernel_spectra32(d)
        for t in (threads, 1):
            rows[f"{name}_t{t}"] = cpu_rate(O, d, s, spectra, t, budget if t > 1 else budget / 2)
            log(f"cpu {name}^2 on {t} thread(s): {rows[f'{name}_t{t}']['value']:.3e} cell-updates/s")
    O.set_threads(threads)
    head = rows[f"4096_t{threads}"]
    return {"value": head["value"], "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{head['runs']} run(s) of {head['batch_steps']} step(s) of the same {d4096.width}x"
                      f"{d4096.height} world on {threads} OpenMP threads (C restatement of the reference FP32 "
                      f"Stockham step, median of runs, {head['seconds']:.1f} s)",
            "rows": rows, "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count()}


def one_step_error(d, sim):
    """One GPU step vs the oracle's FP64 step from the same input state."""
    import numpy as np
    from oracle import oracle as O
    s = sim.read_field("state")
    sim.step(1)
    got = sim.read_field("state")
    want = O.step64(d, s, "fft")
    diff = got.astype(np.float64) - want
    norm = float(np.linalg.norm(want))
    return s, got, {"max_abs": float(np.abs(diff).max()),
                    "rel_l2": float(np.linalg.norm(diff)) / norm if norm > 0 else None,
                    "live_cells": int(np.count_nonzero(want))}


def price_fft(r, w, h, copy_gbps):
    kern = r["kernels"]
    for name, k in kern.items():
        if name in FFT_KERNELS:
            k["bytes"] = kernel_bytes(name, w, h)
            k["GBps"] = k["bytes"] / (k["avg_us"] * 1e-6) / 1e9
    priced = [n for n in kern if "bytes" in kern[n] and kern[n]["launches_per_step"] >= 0.5]
    dom = max(priced, key=lambda n: kern[n]["avg_us"] * kern[n]["launches_per_step"])
    step_bytes = sum(kern[n]["bytes"] * kern[n]["launches_per_step"] for n in priced)
    step_us = sum(kern[n]["avg_us"] * kern[n]["launches_per_step"] for n in priced)
    traffic = RECORDED_TRAFFIC.get((dom, w, h))
    return {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["GBps"], "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": kern[dom]["GBps"] / HBM_PEAK_GBPS, "traffic": traffic,
            "traffic_source": TRAFFIC_SOURCE if traffic else None,
            "bytes_per_launch": kern[dom]["bytes"], "avg_us": kern[dom]["avg_us"],
            "samples": kern[dom]["timed_launches"],
            "step": {"bytes": step_bytes, "kernel_us": step_us, "wall_us": 1e3 * r["ms_per_step"],
                     "achieved": step_bytes / (step_us * 1e-6) / 1e9,
                     "frac": step_bytes / (step_us * 1e-6) / 1e9 / HBM_PEAK_GBPS,
                     "wall_frac": step_bytes / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBPS},
            "copy_probe_GBps": copy_gbps}


def price_direct(r, d, sim):
    w, h = d.width, d.height
    kern = r["kernels"]
    name = "direct_bf16_update"
    fpc = dense_box_flops_per_cell(d)
    tflops = fpc * w * h / (kern[name]["avg_us"] * 1e-6) / 1e12
    kern[name]["flops"] = fpc * w * h
    kern[name]["TFLOPs"] = tflops
    traffic = RECORDED_TRAFFIC.get((name, w, h))
    info = sim.info()
    return {"bound": "mfma", "kernel": name, "achieved": tflops, "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": tflops / BF16_PEAK_TFLOPS, "traffic": traffic,
            "traffic_source": TRAFFIC_SOURCE if traffic else None,
            "flops_per_cell": fpc, "flops_basis": "dense (2*ceil(R)+1)^2 box, 2 flops per tap",
            "useful_flops_per_cell": useful_flops_per_cell(d),
            "avg_us": kern[name]["avg_us"], "samples": kern[name]["timed_launches"],
            "direct_tile": list(sim.direct_tile()),
            "step": {"kernel_us": sum(k["avg_us"] * k["launches_per_step"] for k in kern.values()),
                     "wall_us": 1e3 * r["ms_per_step"]}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="steps per timed block (blocks repeat for >= 1 s)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--path", default="fft_fp32", choices=["fft_fp32", "direct_bf16"],
                    help="path whose rate is `value` (both are timed)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 384-world batch")
    ap.add_argument("--paths-only", action="store_true",
                    help="time the two single-world paths only (for counter runs under a profiler)")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned surrogate section")
    ap.add_argument("--surrogate", action="store_true", help=argparse.SUPPRESS)  # on by default now
    ap.add_argument("--surrogate-train-steps", type=int, default=1000)
    ap.add_argument("--surrogate-radius", type=float, default=0.0,
                    help="with --surrogate-only: train and judge on a 1024^2 one-ring world of this radius instead "
                         "of the 4096^2 three-ring world")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="run only the surrogate section and print it (for profiler runs)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the 384-world batch: 100 warm-up steps, then --steps steps in one timed block "
                         "(for profiler runs)")
    ap.add_argument("--cpu-seconds", type=float, default=4.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and FP64 checks")
    ap.add_argument("--torch-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--torch-surrogate", action="store_true",
                    help="also train the surrogate's network in PyTorch (MIOpen) for comparison (minutes of "
                         "MIOpen algorithm search on a fresh machine)")
    ap.add_argument("--torch-surrogate-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--torch-conv", action="store_true",
                    help="also time the direct path's 53x53 convolution as torch conv2d (MIOpen) for comparison "
                         "(minutes of MIOpen algorithm search on a fresh machine)")
    ap.add_argument("--torch-conv-leg", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.torch_leg:
        return torch_leg(args.torch_leg)
    if args.torch_surrogate_leg:
        return torch_surrogate_leg(args.torch_surrogate_leg)
    if args.torch_conv_leg:
        return torch_conv_leg(args.torch_conv_leg)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this engine has no multi-GPU path (asked for {args.gpus}). "
              "A batch of worlds shards by world and one large world by slab FFT, but neither is built.",
              file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))

    if args.surrogate_only:
        sur = time_surrogate(args.seed, args.surrogate_train_steps, radius=args.surrogate_radius)
        if args.torch_surrogate:
            sur["torch_same_network"] = time_torch_surrogate()
            sur["torch_over_engine_train_ms"] = sur["torch_same_network"]["train_ms_per_step"] / sur["train_ms_per_step"]
        print(json.dumps({"surrogate": sur}), flush=True)
        return 0
    if args.batch_only:
        print(json.dumps({"batch": batch_profile_run(args.seed, args.steps)}), flush=True)
        return 0
    from auto_amd import _abi
    import ctypes
    lib = _abi.load()
    probe_bytes = 1 << 28
    ms = ctypes.c_double()
    _abi.check(lib.ah_copy_probe(probe_bytes, 20, ctypes.byref(ms)))
    copy_gbps = 2 * probe_bytes / (ms.value * 1e-3) / 1e9

    results = {}
    sims = {}
    for path in ("fft_fp32", "direct_bf16"):
        d, sim, r = time_path(path, args.side, args.steps, args.warmup, args.seed)
        log(f"{path}: {r['value']:.3e} cell-updates/s, {r['ms_per_step']:.4f} ms/step (median of {r['blocks']} "
            f"blocks of {args.steps}; p10-p90 {r['p10']:.3e}-{r['p90']:.3e})")
        if path == "fft_fp32":
            roof = price_fft(r, d.width, d.height, copy_gbps)
        else:
            roof = price_direct(r, d, sim)
        results[path] = dict(r, roofline=roof)
        sims[path] = (d, sim)

    head = results[args.path]
    d = sims[args.path][0]
    out = {
        "metric": "Lenia cell-updates/s at 4096^2 (plain FP32 FFT vs BF16 MFMA) at <=1e-3 err",
        "value": head["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.path == "fft_fp32" else "bf16",
        "data": "synthetic",
        "config": {"workload": workload_name(d), "path": args.path, "seed": args.seed},
        "roofline": head["roofline"],
        "timing": {"method": f"median of {head['blocks']} blocks of exactly {args.steps} steps (>= {TIME_S} s, "
                             f">= {MIN_BLOCKS} blocks) after >= {WARMUP_S} s of warm-up; kernel times from "
                             f"HIP events around every {EVENT_EVERY}th launch over a separate "
                             f"{EVENT_EVERY * EVENT_SAMPLES}-step window",
                   "blocks": head["blocks"], "p10": head["p10"], "p90": head["p90"],
                   "reference_method": head["reference_method"]},
        "kernels": head["kernels"],
        "paths": {p: {"value": v["value"], "ms_per_step": v["ms_per_step"], "p10": v["p10"], "p90": v["p90"],
                      "reference_method": v["reference_method"], "roofline": v["roofline"],
                      "kernels": v["kernels"]}
                  for p, v in results.items()},
        "direct_over_fft": results["direct_bf16"]["value"] / results["fft_fp32"]["value"],
    }
    if args.torch_conv:
        tc = time_torch_conv(args.side)
        tc["direct_step_over_conv2d_bf16"] = tc["bf16"]["ms_per_conv"] / results["direct_bf16"]["ms_per_step"]
        out["paths"]["direct_bf16"]["torch_conv2d"] = tc
        log(f"torch conv2d {tc['kernel']}^2 at {tc['side']}^2: bf16 {tc['bf16']['ms_per_conv']:.2f} ms, fp32 "
            f"{tc['fp32']['ms_per_conv']:.2f} ms per convolution")
    if args.paths_only:
        out["cpu_baseline"] = None
        print(json.dumps(out), flush=True)
        return 0
    tf = time_torch_fft(args.side, 200, args.seed)
    log(f"torch rfft2/irfft2 step: {tf['value']:.3e} cell-updates/s")
    tf["engine_fft_over_torch"] = results["fft_fp32"]["value"] / tf["value"]
    out["torch_rocfft"] = tf
    if not args.no_batch:
        out["batch"] = time_batch(args.seed)
        log(f"batch (per-world path, {out['batch']['fft_worlds']} FFT worlds): {out['batch']['value']:.3e} "
            f"cell-updates/s; worst per-world max-abs vs the plain step "
            f"{out['batch']['per_world_vs_plain_step']['worst_max_abs']:.2e}")
        pw = time_plain_worlds(args.seed)
        out["batch"]["plain_per_world"] = pw
        out["batch"]["over_plain_per_world"] = out["batch"]["value"] / pw["value"]
        log(f"the same worlds as {pw['worlds']} standalone FFT Sims: {pw['value']:.3e} cell-updates/s; "
            f"batch / plain per-world stepping = {out['batch']['over_plain_per_world']:.2f}x")
        out["batch_all_direct"] = time_batch(args.seed, path="direct_bf16", check=False)
        log(f"batch (all direct): {out['batch_all_direct']['value']:.3e} cell-updates/s")
    mc = {}
    for path in ("fft_fp32", "direct_bf16"):
        dm, sm, rm = time_multichannel(path, 100, args.seed)
        log(f"multichannel {path}: {rm['value']:.3e} cell-updates/s")
        mc[path] = rm
    out["multichannel"] = mc
    rules = {}
    for rule in ("smoothlife", "mnca"):
        rules[rule] = time_rule(rule, args.side, 100, args.seed)
        log(f"{rule}: {rules[rule]['value']:.3e} cell-updates/s")
    out["rules"] = rules
    if not args.no_surrogate:
        out["surrogate"] = time_surrogate(args.seed, args.surrogate_train_steps)
        log(f"surrogate: {out['surrogate']['value']:.3e} cell-updates/s judged, "
            f"{out['surrogate']['advance']['fallbacks']} fallback(s)")
        if args.torch_surrogate:
            ts = time_torch_surrogate()
            out["surrogate"]["torch_same_network"] = ts
            out["surrogate"]["torch_over_engine_train_ms"] = ts["train_ms_per_step"] / out["surrogate"]["train_ms_per_step"]
            log(f"surrogate network in PyTorch: {ts['train_ms_per_step']:.2f} ms per Adam step")
    if not args.no_cpu and rank == 0:
        log("FP64 checks along the runs")
        out["run_4096"] = run_4096(args.seed)
        log(f"mass drift over 1000 steps at 4096^2, direct vs FFT: {out['run_4096']['mass_drift_direct_vs_fft']:.2e}")
        out["multichannel"]["errors"] = multichannel_errors(args.seed)
        log("cpu baseline")
        dd, sim = sims["fft_fp32"]
        s0, _, err = one_step_error(dd, sim)
        out["err_vs_fp64_one_step"] = {"fft_fp32": err}
        _, _, out["err_vs_fp64_one_step"]["direct_bf16"] = one_step_error(*sims["direct_bf16"])
        out["cpu_baseline"] = cpu_baseline(dd, s0, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
