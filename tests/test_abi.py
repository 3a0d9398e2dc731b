"""CPU checks of the C ABI: the library loads, exports every declared
symbol, and the host-side logic (registry, validation) behaves without a GPU."""
import ctypes
import os
import re

import pytest

from auto_amd import _abi, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "auto_hip.h")).read()
    return sorted(set(re.findall(r"\b(ah_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.symbols())


def test_struct_sizes_match_header(tmp_path):
    """ctypes layouts equal the C compiler's sizeof for every ABI struct."""
    import shutil
    import subprocess
    names = ["ah_kernel", "ah_lenia_desc", "ah_stats", "ah_info", "ah_kernel_time", "ah_direct_tile",
             "ah_batch_tune", "ah_batch_group_tune", "ah_batch_info", "ah_surrogate_report"]
    types = [_abi.Kernel, _abi.LeniaDesc, _abi.Stats, _abi.Info, _abi.KernelTime, _abi.DirectTile,
             _abi.BatchTune, _abi.BatchGroupTune, _abi.BatchInfo, _abi.SurrogateReport]
    want = [52, 508, 64, 56, 56, 12, 24, 24, 1800, 64]  # x86-64 gcc, when no compiler is on the path
    gcc = shutil.which("gcc")
    if gcc:
        src = tmp_path / "sz.c"
        fmt = " ".join(["%zu"] * len(names))
        sizes = ",".join(f"sizeof({n})" for n in names)
        src.write_text(f'#include "include/auto_hip.h"\n#include <stdio.h>\n'
                       f'int main(void){{printf("{fmt}",{sizes});return 0;}}\n')
        exe = tmp_path / "sz"
        subprocess.run([gcc, f"-I{ROOT}", str(src), "-o", str(exe)], check=True)
        want = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    got = [ctypes.sizeof(t) for t in types]
    assert got == want


def test_registry_matches_reference_entry():
    """ah_lenia_build("lenia") == the reference entry (spectral.rs:286-352,482-490),
    sizes rounded up to powers of two (mod.rs:51-61)."""
    lib = _abi.load()
    d = _abi.LeniaDesc()
    assert lib.ah_lenia_build(b"lenia", 100, 60, ctypes.byref(d)) == 0
    ref = W.lenia_reference(100, 60)
    assert (d.width, d.height) == (128, 64) == (ref.width, ref.height)
    assert bytes(d) == bytes(ref)
    assert lib.ah_lenia_build(b"life", 64, 64, ctypes.byref(d)) == _abi.ERR_INVALID
    assert b"unknown" in lib.ah_last_error()


def test_registry_smoothlife_and_mnca():
    """"smoothlife" and "mnca" as the reference builds them (spectral.rs:354-510),
    and "mnca.i0,o0,i1,o1" custom rings (mod.rs:86-92)."""
    lib = _abi.load()
    d = _abi.LeniaDesc()
    assert lib.ah_lenia_build(b"smoothlife", 300, 200, ctypes.byref(d)) == 0
    assert bytes(d) == bytes(W.smoothlife(300, 200)) and (d.width, d.height) == (512, 256)
    assert (d.kernels[0].radius, d.kernels[1].inner, d.kernels[1].radius) == (7.0, 7.0, 21.0)
    assert lib.ah_lenia_build(b"mnca", 256, 256, ctypes.byref(d)) == 0
    assert bytes(d) == bytes(W.mnca(256, 256))
    assert lib.ah_lenia_build(b"mnca.10,15.5,2,6", 256, 256, ctypes.byref(d)) == 0
    assert bytes(d) == bytes(W.mnca(256, 256, rings=(10, 15.5, 2, 6)))
    for bad in (b"mnca.1,2", b"mnca.1,2,3,x", b"mnca.", b"mnca.1,2,3,4,5"):
        assert lib.ah_lenia_build(bad, 64, 64, ctypes.byref(d)) == _abi.ERR_INVALID, bad


@pytest.mark.parametrize("mutate,code", [
    (lambda d: setattr(d, "rule", 5), _abi.ERR_INVALID),
    (lambda d: setattr(d, "n_kernels", 1), _abi.ERR_INVALID),
    (lambda d: setattr(d, "path", _abi.PATH_DIRECT_BF16), _abi.ERR_UNSUPPORTED),
    (lambda d: setattr(d.kernels[0], "shape", 9), _abi.ERR_INVALID),
    (lambda d: setattr(d.kernels[1], "aa", 0.0), _abi.ERR_INVALID),
    (lambda d: setattr(d.kernels[1], "inner", 30.0), _abi.ERR_INVALID),
    (lambda d: setattr(d.kernels[1], "radius", 31.5), _abi.ERR_INVALID),  # reach 32: world smaller than its kernel
])
def test_invalid_smoothlife_descriptors(mutate, code):
    lib = _abi.load()
    d = W.smoothlife(64, 64)
    mutate(d)
    h = ctypes.c_void_p()
    assert lib.ah_sim_new(ctypes.byref(d), 1, ctypes.byref(h)) == code
    assert not h.value


@pytest.mark.parametrize("mutate,code", [
    (lambda d: setattr(d, "width", 100), _abi.ERR_INVALID),
    (lambda d: setattr(d.kernels[0], "radius", 40.0), _abi.ERR_INVALID),  # world smaller than its kernel
    (lambda d: setattr(d, "n_kernels", 0), _abi.ERR_INVALID),
    (lambda d: setattr(d.kernels[0], "sigma", 0.0), _abi.ERR_INVALID),
    (lambda d: setattr(d.kernels[0], "src", 2), _abi.ERR_INVALID),
    (lambda d: setattr(d, "path", 7), _abi.ERR_INVALID),
    (lambda d: setattr(d, "dt", float("nan")), _abi.ERR_INVALID),
])
def test_invalid_descriptors_are_errors_not_crashes(mutate, code):
    """The reference panics on these (spectral.rs:103-116); the ABI returns an error."""
    lib = _abi.load()
    d = W.lenia_reference(64, 64)
    mutate(d)
    h = ctypes.c_void_p()
    assert lib.ah_sim_new(ctypes.byref(d), 1, ctypes.byref(h)) == code
    assert not h.value
    assert lib.ah_last_error()


def test_null_arguments():
    lib = _abi.load()
    assert lib.ah_sim_step(None, 1, None) == _abi.ERR_INVALID
    assert lib.ah_sim_wait(None) == _abi.ERR_INVALID
    lib.ah_sim_destroy(None)


def test_workload_shapes():
    d = W.rings3_4096()
    assert (d.width, d.height, d.n_kernels, d.kernels[0].rings) == (4096, 4096, 1, 3)
    assert list(d.kernels[0].beta) == pytest.approx([1, 2 / 3, 1 / 3])
    m = W.multichannel_1024(seed=1)
    assert m.channels == 3 and m.n_kernels == 9
    for k in range(9):
        kd = m.kernels[k]
        assert (kd.src, kd.dst) == (k // 3, k % 3)
        assert 8 <= kd.radius <= 20 and 1 <= kd.rings <= 3
        assert 0.1 <= kd.mu <= 0.35 and 0.01 <= kd.sigma <= 0.05 and 0.2 <= kd.weight <= 1
    ws = W.uneven_worlds(seed=1)
    assert len(ws) == 384
    assert {d.width for d, _ in ws} <= {128, 256, 512, 1024}
