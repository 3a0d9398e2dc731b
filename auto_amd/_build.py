"""Build libauto_hip.so in-tree with hipcc for gfx950.

Each source compiles to its own object in parallel (auto_amd/_build/), then
one link makes the shared library.  Usage: python -m auto_amd._build (or
__graft_entry__.build()).  `out` and `defines` let a caller build a variant
library elsewhere (AUTO_HIP_LIB selects which one the package loads).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["engine.cpp", "init.hip", "fft_step.hip", "direct_bf16.hip", "batch.cpp", "surrogate.hip", "surrogate.cpp", "render.hip"]
OUT = os.path.join(HERE, "libauto_hip.so")
OBJ_DIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay off for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
NOISE = ("unknown target feature", "packed-fp32-ops")


def _clean(stderr: str) -> str:
    return "\n".join(l for l in stderr.splitlines() if not any(n in l for n in NOISE))


# Host-side AddressSanitizer (device code is never instrumented): for the
# CPU sanitizer test, which loads such a build with the sanitizer runtime
# preloaded.
ASAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]


def _compile(src: str, obj: str, defines, extra=()) -> tuple[int, str, list]:
    cmd = [HIPCC, *FLAGS, *extra, *[f"-D{d}" for d in defines], "-c", "-o", obj, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, _clean(r.stderr), cmd


def build(verbose: bool = False, out: str = OUT, defines=(), obj_dir: str = OBJ_DIR, asan: bool = False) -> str:
    extra = ASAN_HOST if asan else []
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(obj_dir, s + ".o") for s in SOURCES]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda so: _compile(so[0], so[1], defines, extra), zip(srcs, objs)))
    for (rc, err, cmd), src in zip(results, srcs):
        if verbose:
            print(" ".join(cmd), flush=True)
        if rc != 0:
            raise RuntimeError(f"hipcc failed on {os.path.basename(src)} ({rc}):\n{err}")
        if verbose and err.strip():
            print(err, file=sys.stderr)
    tmp = out + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *extra, "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed ({r.returncode}):\n{_clean(r.stderr)}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
