"""ah_state_stats_noalloc: the stats reduction over a caller-held device
state with its partials in the library's fixed scratch array, allocating
nothing.  Checked against ah_sim_stats on the same state (bit for bit: the
same kernels and block count), on two calls in a row with the second one
smaller (fewer cells and fewer channels), so stale partials of the first
call would show.  The device buffers are the test's own, made through the
HIP runtime the engine itself links (libamdhip64.so.7)."""
import ctypes

import numpy as np
import pytest

from auto_amd import Sim, workloads as W
from auto_amd._abi import Stats, load

pytestmark = pytest.mark.gpu

H2D = 1  # hipMemcpyHostToDevice


def _hip():
    load()  # the engine first, so its HIP runtime is the one this handle names
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _device_copy(hip, a):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _noalloc(lib, p, cells, channels):
    st = Stats()
    rc = lib.ah_state_stats_noalloc(p, cells, channels, None, ctypes.byref(st))
    assert rc == 0, lib.ah_last_error()
    return st


def _sim_stats(d, state):
    sim = Sim(d, 1)
    sim.write_field("state", state)
    return sim.stats()


def _same(a, b, channels):
    for c in range(channels):
        assert a.mass[c] == b.mass[c], (c, a.mass[c], b.mass[c])
        assert a.min[c] == b.min[c] and a.max[c] == b.max[c], c
    assert a.nonfinite == b.nonfinite


def test_noalloc_stats_equal_sim_stats_on_two_calls():
    lib = load()
    hip = _hip()
    rng = np.random.default_rng(11)
    # First call: 3 channels of 2048^2 (256 partials per channel), with a
    # few non-finite cells.
    d1 = W.multichannel_1024(seed=1, side=2048)
    s1 = rng.random((3, 2048, 2048), dtype=np.float32)
    s1[1, 5, 7] = np.nan
    s1[2, 100, 3] = np.inf
    # Second call: 1 channel of 256^2 (4 partials), smaller in both.
    d2 = W.single_ring_256()
    s2 = (rng.random((1, 256, 256), dtype=np.float32) * 0.5).astype(np.float32)
    p1, p2 = _device_copy(hip, s1), _device_copy(hip, s2)
    try:
        got1 = _noalloc(lib, p1, 2048 * 2048, 3)
        got2 = _noalloc(lib, p2, 256 * 256, 1)
    finally:
        hip.hipFree(p1)
        hip.hipFree(p2)
    want1, want2 = _sim_stats(d1, s1), _sim_stats(d2, s2)
    _same(got1, want1, 3)
    _same(got2, want2, 1)
    assert got1.nonfinite == 2 and got2.nonfinite == 0
    m = float(s2.sum(dtype=np.float64))
    assert abs(got2.mass[0] - m) <= 1e-9 * m
    print(f"noalloc stats: mass {got1.mass[0]:.6f}/{got1.mass[1]:.6f}/{got1.mass[2]:.6f}, then {got2.mass[0]:.6f}")


def test_noalloc_stats_rejects_bad_sizes():
    lib = load()
    st = Stats()
    fake = ctypes.c_void_p(256)  # never dereferenced: every call below fails its checks first
    assert lib.ah_state_stats_noalloc(fake, 1024, 4, None, ctypes.byref(st)) == 1
    assert b"channels" in lib.ah_last_error()
    assert lib.ah_state_stats_noalloc(fake, 1022, 1, None, ctypes.byref(st)) == 1
    assert lib.ah_state_stats_noalloc(fake, 0, 1, None, ctypes.byref(st)) == 1
    assert lib.ah_state_stats_noalloc(ctypes.c_void_p(260), 1024, 1, None, ctypes.byref(st)) == 1
    assert lib.ah_state_stats_noalloc(None, 1024, 1, None, ctypes.byref(st)) == 1
