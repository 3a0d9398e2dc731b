"""auto_amd.bench_rows: rows in the reference's results format (GPU)."""
import json

import pytest

from auto_amd import bench_rows

# The fields of `Sample`, crates/auto-cli/src/bench.rs:39-54, in order.
SAMPLE_FIELDS = ["id", "tag", "adapter", "checks", "width", "height", "cells", "memory_bytes", "steps_per_second",
                 "generations_per_second", "cells_per_second", "batch_steps", "batches", "spread"]


@pytest.mark.gpu
def test_rows_and_verify(tmp_path, capsys):
    out = tmp_path / "gpu.jsonl"
    for id_ in ("lenia", "smoothlife", "mnca"):
        rc = bench_rows.main([id_, "--sizes", "256x256,128x128", "--seconds", "0.2", "--warmup", "4",
                              "--out", str(out), "--tag", "test", "--verify"])
        assert rc == 0
    err = capsys.readouterr().err
    assert err.count("verified:") == 6 and "MISMATCH" not in err
    rows = [json.loads(l) for l in out.read_text().splitlines()]
    assert len(rows) == 6
    for r in rows:
        assert list(r) == SAMPLE_FIELDS
        assert r["cells"] == r["width"] * r["height"] and r["cells_per_second"] > 0
        assert "gfx950" in r["adapter"] and r["batches"] >= 3
    assert [(r["id"], r["width"]) for r in rows[:2]] == [("lenia", 128), ("lenia", 256)]  # small to big
    assert bench_rows.main(["life", "--sizes", "64x64", "--out", str(out)]) == 2


@pytest.mark.gpu
def test_row_at_8192(tmp_path):
    """A Lenia row at 8192^2, the largest size the reference's bench runs
    (scripts/bench_all.sh:50-52, results/gpu.jsonl:82)."""
    out = tmp_path / "gpu.jsonl"
    assert bench_rows.main(["lenia", "--sizes", "8192x8192", "--seconds", "0.3", "--warmup", "2",
                            "--out", str(out), "--tag", "test"]) == 0
    (r,) = [json.loads(l) for l in out.read_text().splitlines()]
    assert (r["width"], r["height"]) == (8192, 8192) and r["cells_per_second"] > 0
    print(f"lenia 8192^2: {r['cells_per_second']:.3e} cell-updates/s")
