"""Writes tests/golden/pcg_vectors.npz: known answers for the reference's
per-cell random source, taken from the reference's own WGSL text.

The reference defines `pcg`, `rand3` and `unit_f` in the WGSL prelude every
automaton shares (crates/auto-core/src/program.rs:139-146).  This script reads
that file as text, pulls out the constants and shift amounts of `pcg`, the
nesting of `rand3` and the scale of `unit_f` with patterns that must match the
reference's exact shape (the script fails otherwise), and evaluates them in
numpy with 32-bit wrapping arithmetic.  Nothing is typed in by hand, so a
constant mistyped in the oracle's C restatement fails against these vectors.

Run from the repo root, where the reference is checked out at
/root/reference: python tests/golden/make_pcg_vectors.py
"""
import os
import re

import numpy as np

SRC = "/root/reference/crates/auto-core/src/program.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_vectors.npz")


def parse(text: str):
    m = re.search(r"fn pcg\(v: u32\) -> u32 \{\s*"
                  r"let s = v \* (\d+)u \+ (\d+)u;\s*"
                  r"let w = \(\(s >> \(\(s >> (\d+)u\) \+ (\d+)u\)\) \^ s\) \* (\d+)u;\s*"
                  r"return \(w >> (\d+)u\) \^ w;\s*\}", text)
    if not m:
        raise SystemExit("pcg: the reference's definition no longer has the expected shape")
    mul, inc, sh1, add, mul2, sh2 = (int(g) for g in m.groups())
    if not re.search(r"fn rand3\(a: u32, b: u32, c: u32\) -> u32 \{ return pcg\(pcg\(pcg\(a\) \^ b\) \^ c\); \}", text):
        raise SystemExit("rand3: the reference's definition no longer has the expected shape")
    u = re.search(r"fn unit_f\(r: u32\) -> f32 \{ return f32\(r\) \* \(1\.0 / (\d+)\.0\); \}", text)
    if not u:
        raise SystemExit("unit_f: the reference's definition no longer has the expected shape")
    return dict(mul=mul, inc=inc, sh1=sh1, add=add, mul2=mul2, sh2=sh2, denom=int(u.group(1)))


def evaluate(p, v, a, b, c, r):
    M = np.uint64(0xFFFFFFFF)

    def pcg(x):
        x = x.astype(np.uint64)
        s = (x * np.uint64(p["mul"]) + np.uint64(p["inc"])) & M
        w = (((s >> ((s >> np.uint64(p["sh1"])) + np.uint64(p["add"]))) ^ s) * np.uint64(p["mul2"])) & M
        return ((w >> np.uint64(p["sh2"])) ^ w) & M

    out_pcg = pcg(v).astype(np.uint32)
    out_rand3 = pcg(pcg(pcg(a) ^ b.astype(np.uint64)) ^ c.astype(np.uint64)).astype(np.uint32)
    # f32(r) rounds to nearest even (exact in float64 first); the scale is a power of two.
    scale = np.float32(1.0 / p["denom"])
    out_unit = (r.astype(np.float64).astype(np.float32) * scale).astype(np.float32)
    return out_pcg, out_rand3, out_unit


def main():
    with open(SRC) as f:
        p = parse(f.read())
    rng = np.random.default_rng(20261204)
    edge = np.array([0, 1, 2, 3, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF, 0x12345678, 0xDEADBEEF],
                    np.uint32)
    v = np.concatenate([edge, rng.integers(0, 1 << 32, 1014, dtype=np.uint64).astype(np.uint32)])
    # rand3 as the seeds call it: (x, y, seed) over a small grid and random triples.
    gx, gy = np.meshgrid(np.arange(64, dtype=np.uint32), np.arange(16, dtype=np.uint32))
    a = np.concatenate([gx.ravel(), rng.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)])
    b = np.concatenate([gy.ravel(), rng.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)])
    c = np.concatenate([np.full(64 * 16, 5, np.uint32), rng.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)])
    r = np.concatenate([edge, np.array([0xFFFFFF7F, 0xFFFFFF80, 0xFFFFFF81, 0x00FFFFFF, 0x01000001], np.uint32),
                        rng.integers(0, 1 << 32, 1009, dtype=np.uint64).astype(np.uint32)])
    out_pcg, out_rand3, out_unit = evaluate(p, v, a, b, c, r)
    np.savez_compressed(OUT, v=v, pcg=out_pcg, a=a, b=b, c=c, rand3=out_rand3, r=r, unit_f=out_unit)
    print(f"wrote {OUT}: {len(v)} pcg, {len(a)} rand3, {len(r)} unit_f vectors from {SRC}")


if __name__ == "__main__":
    main()
