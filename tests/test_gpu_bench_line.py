"""bench.py's output line: one JSON object on stdout with the fields and
units the benchmark contract names (a short run at a small size)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                       timeout=240, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    return json.loads(lines[0])


def test_bench_prints_one_contract_line():
    b = _run("--steps", "5", "--warmup", "2", "--side", "1024", "--paths-only", "--no-cpu")
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline"):
        assert key in b, key
    assert b["n_gpus"] == 1 and b["steps"] == 5 and b["warmup"] == 2
    assert b["unit"] == "cell-updates/s" and b["higher_is_better"] is True and b["dtype"] == "f32"
    assert "workload" in b["config"]
    assert b["value"] > 0 and abs(b["value"] - 1024 * 1024 / (b["ms_per_step"] * 1e-3)) / b["value"] < 1e-6
    r = b["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert 0 < r["frac"] < 1 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-9
    assert r["traffic"] is None  # recorded only for the 4096^2 world
    # Median of repeated blocks of exactly --steps steps, >= 256 event
    # samples behind the roofline's kernel time, and the per-step kernel sum
    # within the measured wall time per step, give or take the ~1.5 us of
    # dispatch an event pair adds to each timed launch (2 launches per step).
    assert b["timing"]["blocks"] >= 3 and b["timing"]["p10"] <= b["value"] * 1.0001
    assert r["samples"] >= 256
    assert r["step"]["kernel_us"] <= r["step"]["wall_us"] * 1.02 + 2 * 2.0
    dr = b["paths"]["direct_bf16"]["roofline"]
    assert dr["bound"] == "mfma" and dr["flops_per_cell"] == 2.0 * 53 ** 2 and dr["samples"] >= 256
