"""Benchmark rows in the reference's results format, and a verify mode.

    python -m auto_amd.bench_rows lenia --sizes 1024x1024,4096x4096 \\
        --seconds 3 --out results/gpu.jsonl --tag hip [--verify]

Measures like `auto-cli bench` (crates/auto-cli/src/bench.rs:112-165): run
`--warmup` steps, grow the batch until one batch takes >= 20 ms, then time
batches for `--seconds` and keep the median rate and the p90-p10 spread.
Each size appends one JSON line with the fields of the reference's `Sample`
(bench.rs:39-54), which scripts/make_report.py reads from results/*.jsonl
(make_report.py:202-214).

--verify plays the part of auto-vk's verify (crates/auto-vk/src/main.rs:
658-763), which compares two backends after 10 steps: for Lenia-rule ids it
runs 10 steps on the FFT FP32 path and on the independent direct BF16
three-term path from the same seed and reports their max-abs difference
(bitwise equality is not expected between two convolution algorithms).
SmoothLife and MNCA run on the FFT path only, so they report the
split-call check instead: 10 steps in one call equal 10 single steps with
the state written back in between.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _abi
from .sim import Sim, build


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    _abi.check(_abi.load().ah_device_name(buf, len(buf)))
    return buf.value.decode()


def measure(desc, seconds: float, warmup: int) -> dict:
    sim = Sim(desc, 1)
    sim.step(warmup)
    sim.wait()
    batch = 1
    while True:
        t = time.perf_counter()
        sim.step(batch)
        sim.wait()
        dt = time.perf_counter() - t
        if dt >= 0.02 or batch >= 1 << 16:
            break
        want = int(np.ceil(batch * 0.025 / max(dt, 1e-5)))
        batch = min(max(want, batch + 1), batch * 8, 1 << 16)
    rates = []
    start = time.perf_counter()
    while time.perf_counter() - start < seconds or len(rates) < 3:
        t = time.perf_counter()
        sim.step(batch)
        sim.wait()
        rates.append(batch / (time.perf_counter() - t))
    rates.sort()
    median = rates[len(rates) // 2]
    spread = (rates[len(rates) * 9 // 10] - rates[len(rates) // 10]) / median
    info = sim.info()
    gens = info.generations_per_step
    return {"cells": info.cells_per_step, "memory_bytes": info.total_bytes, "steps_per_second": median,
            "generations_per_second": median * gens, "cells_per_second": median * gens * info.cells_per_step,
            "batch_steps": batch, "batches": len(rates), "spread": spread}


def verify(id_: str, w: int, h: int) -> str:
    d = build(id_, w, h)
    if d.rule == _abi.RULE_LENIA:
        a = Sim(d, 1)
        d3 = build(id_, w, h)
        d3.path = _abi.PATH_DIRECT_BF16_X3
        b = Sim(d3, 1)
        a.step(10)
        b.step(10)
        err = float(np.abs(a.read_field("state") - b.read_field("state")).max())
        ok = err < 1e-3
        return f"{'verified' if ok else 'MISMATCH'}: FFT FP32 and direct BF16x3 paths differ by {err:.2e} max-abs " \
               f"after 10 steps of {id_} {d.width}x{d.height}"
    a, b = Sim(d, 1), Sim(d, 1)
    a.step(10)
    for _ in range(10):
        b.step(1)
        b.write_field("state", b.read_field("state"))
    ok = np.array_equal(a.read_field("state"), b.read_field("state"))
    return f"{'verified' if ok else 'MISMATCH'}: 10 steps in one call and 10 resumed single steps give " \
           f"{'identical' if ok else 'different'} bytes for {id_} {d.width}x{d.height}"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("id")
    ap.add_argument("--sizes", default="1024x1024")
    ap.add_argument("--seconds", type=float, default=3.0)
    ap.add_argument("--out", default="results/bench.jsonl")
    ap.add_argument("--tag", default="")
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--path", default="fft_fp32", choices=sorted(_abi.PATHS))
    ap.add_argument("--verify", action="store_true")
    args = ap.parse_args(argv)
    sizes = sorted((tuple(int(v) for v in s.split("x")) for s in args.sizes.split(",")), key=lambda s: s[0] * s[1])
    adapter = device_name()
    for w, h in sizes:
        try:
            d = build(args.id, w, h)
        except _abi.EngineError as e:
            print(f"unknown id {args.id} or size {w}x{h} unsupported: {e}", file=sys.stderr)
            return 2
        d.path = _abi.PATHS[args.path]
        if args.verify:
            msg = verify(args.id, w, h)
            print(msg, file=sys.stderr)
            if msg.startswith("MISMATCH"):
                return 1
        try:
            m = measure(d, args.seconds, args.warmup)
        except _abi.EngineError as e:
            print(f"{args.id} {w}x{h}: {e}", file=sys.stderr)
            continue
        row = {"id": args.id, "tag": args.tag, "adapter": adapter, "checks": "none", "width": d.width,
               "height": d.height, **m}
        print(f"{args.id:<26} {d.width:>7}x{d.height:<7} {m['memory_bytes'] / 1e6:>9.1f} MB "
              f"{m['generations_per_second']:>12.1f} gen/s {m['cells_per_second'] / 1e9:>14.3f} Gcell/s  "
              f"batch {m['batch_steps']} spread {m['spread'] * 100:.1f}%")
        if os.path.dirname(args.out):
            os.makedirs(os.path.dirname(args.out), exist_ok=True)
        with open(args.out, "a") as f:
            f.write(json.dumps(row) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
