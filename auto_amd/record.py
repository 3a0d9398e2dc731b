"""Record a run as frames, as `auto-cli record` does (crates/auto-cli/src/
record.rs:51-235) minus its ffmpeg pipe: every frame is rendered on the GPU
(Sim.render) and written as a binary PPM, which ffmpeg or any viewer reads.

    python -m auto_amd.record smoothlife --size 1024x1024 --frames 120 \\
        --steps-per-frame 2 --frame-size 512x512 --out frames/
"""
from __future__ import annotations

import argparse
import os
import sys

from .sim import Sim, build


def write_ppm(path: str, rgba) -> None:
    h, w = rgba.shape[:2]
    with open(path, "wb") as f:
        f.write(f"P6 {w} {h} 255\n".encode())
        f.write(rgba[..., :3].tobytes())


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("id")
    ap.add_argument("--size", default="1024x1024")
    ap.add_argument("--frame-size", default="512x512")
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--steps-per-frame", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default="frames")
    args = ap.parse_args(argv)
    w, h = (int(v) for v in args.size.split("x"))
    fw, fh = (int(v) for v in args.frame_size.split("x"))
    sim = Sim(build(args.id, w, h), args.seed)
    os.makedirs(args.out, exist_ok=True)
    for i in range(args.frames):
        write_ppm(os.path.join(args.out, f"frame_{i:05d}.ppm"), sim.render(fw, fh))
        sim.step(args.steps_per_frame)
    print(f"{args.frames} frames of {args.id} {sim.width}x{sim.height} in {args.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
