"""Host code under AddressSanitizer, on the CPU.

The engine's host side (registry, descriptor validation, the ABI's argument
checks) is rebuilt with AddressSanitizer on the host only (device code is
never instrumented), the oracle with gcc's AddressSanitizer, and the CPU test
files that exercise them run in child processes with the matching sanitizer
runtime preloaded.  A sanitizer report fails the test.  Without a GPU every
device call fails cleanly, so this covers the paths that run before any
device work.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.slow


def _run(env_extra, files):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", **env_extra)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", *files],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return out


def test_engine_host_code_under_asan(tmp_path):
    from auto_amd import _build
    runtimes = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not runtimes:
        pytest.skip("clang AddressSanitizer runtime not found")
    lib = _build.build(out=str(tmp_path / "libauto_hip_asan.so"), obj_dir=str(tmp_path / "obj"), asan=True)
    out = _run({"AUTO_HIP_LIB": lib, "LD_PRELOAD": runtimes[0]}, ["tests/test_abi.py"])
    print(out.strip().splitlines()[-1])


def test_oracle_under_asan():
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(libasan):
        pytest.skip("gcc AddressSanitizer runtime not found")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liblenia_oracle_asan.so"], check=True)
    lib = os.path.join(ROOT, "oracle", "liblenia_oracle_asan.so")
    out = _run({"LENIA_ORACLE_LIB": lib, "LD_PRELOAD": libasan, "OMP_NUM_THREADS": "4"},
               ["tests/test_oracle.py", "tests/test_oracle_rules.py", "tests/test_surrogate_oracle.py"])
    print(out.strip().splitlines()[-1])
