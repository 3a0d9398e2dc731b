"""Python face of the CPU oracle.  TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only to check the engine's results.  The engine never
calls it.

Wraps oracle/liblenia_oracle.so (the C restatement of the reference's Lenia
passes, see lenia_oracle.c) and adds the FP64 FFT convolution, computed with
numpy's FP64 FFT, as a second FP64 method beside the C direct convolution.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from auto_amd._abi import LeniaDesc, Kernel, Stats

HERE = os.path.dirname(os.path.abspath(__file__))
# LENIA_ORACLE_LIB selects another build of the same source (the sanitizer test's).
LIB = os.environ.get("LENIA_ORACLE_LIB", os.path.join(HERE, "liblenia_oracle.so"))

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        sig = {
            "or_pcg": (ctypes.c_uint32, [ctypes.c_uint32]),
            "or_rand3": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "or_unit_f": (ctypes.c_float, [ctypes.c_uint32]),
            "or_splitmix64": (ctypes.c_uint64, [ctypes.POINTER(ctypes.c_uint64)]),
            "or_seed_orbium": (ctypes.c_float, [ctypes.c_uint32] * 3),
            "or_init_state": (ctypes.c_int, [ctypes.POINTER(LeniaDesc), ctypes.c_uint32, _F]),
            "or_core64": (ctypes.c_double, [ctypes.c_double]),
            "or_shell64": (ctypes.c_double, [ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _F]),
            "or_reach": (ctypes.c_int, [ctypes.c_float]),
            "or_kernel_reach": (ctypes.c_int, [ctypes.POINTER(Kernel)]),
            "or_profile64": (ctypes.c_double, [ctypes.POINTER(Kernel), ctypes.c_int, ctypes.c_int]),
            "or_smoothlife32": (ctypes.c_float, [ctypes.c_float] * 4),
            "or_smoothlife64": (ctypes.c_double, [ctypes.c_double] * 4),
            "or_mnca32": (ctypes.c_float, [ctypes.c_float] * 3),
            "or_mnca64": (ctypes.c_double, [ctypes.c_double] * 3),
            "or_kernel_box64": (ctypes.c_double, [ctypes.POINTER(Kernel), _D]),
            "or_kernel_sum_ref32": (ctypes.c_float, [ctypes.c_float]),
            "or_fft_axis_ref32": (None, [ctypes.c_uint32, ctypes.c_uint32, _F, _F, ctypes.c_int, ctypes.c_int]),
            "or_kernel_spectrum_ref32": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _F]),
            "or_kernel_spectrum32": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Kernel), _F]),
            "or_conv_ref32": (None, [ctypes.c_uint32, ctypes.c_uint32, _F, _F, _F]),
            "or_step_ref32": (None, [ctypes.POINTER(LeniaDesc), _F, _F, _F]),
            "or_conv_direct64": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Kernel), _F, _D]),
            "or_update64": (None, [ctypes.POINTER(LeniaDesc), _D, _F, _F]),
            "or_step_direct64": (None, [ctypes.POINTER(LeniaDesc), _F, _F]),
            "or_stats": (None, [_F, ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(Stats)]),
            "or_max_threads": (ctypes.c_int, []),
            "or_set_threads": (None, [ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


# ---------------------------------------------------------------- generators

def pcg(v: int) -> int:
    return lib().or_pcg(v)


def rand3(a: int, b: int, c: int) -> int:
    return lib().or_rand3(a, b, c)


def unit_f(r: int) -> float:
    return lib().or_unit_f(r)


def splitmix64(state: int, n: int) -> list[int]:
    st = ctypes.c_uint64(state)
    return [lib().or_splitmix64(ctypes.byref(st)) for _ in range(n)]


def init_state(desc: LeniaDesc, seed: int) -> np.ndarray:
    out = np.zeros((desc.channels, desc.height, desc.width), np.float32)
    if lib().or_init_state(ctypes.byref(desc), seed, _fp(out)) != 0:
        raise ValueError("bad init in descriptor")
    return out


# ------------------------------------------------------------------- kernels

def kernel_box64(k: Kernel) -> np.ndarray:
    reach = lib().or_kernel_reach(ctypes.byref(k))
    taps = np.zeros((2 * reach + 1, 2 * reach + 1), np.float64)
    lib().or_kernel_box64(ctypes.byref(k), _dp(taps))
    return taps


def kernel_torus64(k: Kernel, w: int, h: int) -> np.ndarray:
    """FP64 normalised kernel on the torus, min-image placement."""
    taps = kernel_box64(k)
    reach = taps.shape[0] // 2
    out = np.zeros((h, w), np.float64)
    for dy in range(-reach, reach + 1):
        for dx in range(-reach, reach + 1):
            out[dy % h, dx % w] += taps[dy + reach, dx + reach]
    return out


def kernel_spectrum_ref32(w: int, h: int, radius: float) -> np.ndarray:
    """The reference init's kernel spectrum (complex64, [h][w])."""
    out = np.zeros((h, w, 2), np.float32)
    lib().or_kernel_spectrum_ref32(w, h, radius, _fp(out))
    return out[..., 0] + 1j * out[..., 1]


def kernel_spectra32(desc: LeniaDesc) -> np.ndarray:
    """Spectra of every kernel through the reference's FP32 passes, interleaved f32."""
    out = np.zeros((desc.n_kernels, desc.height, desc.width, 2), np.float32)
    for k in range(desc.n_kernels):
        lib().or_kernel_spectrum32(desc.width, desc.height, ctypes.byref(desc.kernels[k]), _fp(out[k]))
    return out


# --------------------------------------------------------------------- FFT

def fft_axis_ref32(x: np.ndarray, axis_x: bool, inverse: bool) -> np.ndarray:
    """All Stockham stages of the reference along one axis of a complex64 [h][w] field."""
    h, w = x.shape
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).astype(np.float32))
    tmp = np.empty_like(buf)
    lib().or_fft_axis_ref32(w, h, _fp(buf), _fp(tmp), int(axis_x), int(inverse))
    return buf[..., 0] + 1j * buf[..., 1]


# ------------------------------------------------------------------- steps

def step_ref32(desc: LeniaDesc, s: np.ndarray, spectra: np.ndarray | None = None) -> np.ndarray:
    """One step through the reference's FP32 passes (the CPU reference step)."""
    if spectra is None:
        spectra = kernel_spectra32(desc)
    s = np.ascontiguousarray(s, np.float32)
    out = np.empty_like(s)
    lib().or_step_ref32(ctypes.byref(desc), _fp(spectra), _fp(s), _fp(out))
    return out


def conv_ref32(desc: LeniaDesc, s: np.ndarray, k: int, spectra: np.ndarray | None = None) -> np.ndarray:
    if spectra is None:
        spectra = kernel_spectra32(desc)
    src = np.ascontiguousarray(s.reshape(desc.channels, desc.height, desc.width)[desc.kernels[k].src])
    u = np.empty((desc.height, desc.width), np.float32)
    lib().or_conv_ref32(desc.width, desc.height, _fp(np.ascontiguousarray(spectra[k])), _fp(src), _fp(u))
    return u


def conv_direct64(desc: LeniaDesc, s: np.ndarray, k: int) -> np.ndarray:
    src = np.ascontiguousarray(s.reshape(desc.channels, desc.height, desc.width)[desc.kernels[k].src],
                               np.float32)
    u = np.empty((desc.height, desc.width), np.float64)
    lib().or_conv_direct64(desc.width, desc.height, ctypes.byref(desc.kernels[k]), _fp(src), _dp(u))
    return u


def conv_fft64(desc: LeniaDesc, s: np.ndarray, k: int, khat: np.ndarray | None = None) -> np.ndarray:
    """FP64 periodic convolution by numpy's FP64 real FFT (a second FP64 method)."""
    w, h = desc.width, desc.height
    src = s.reshape(desc.channels, h, w)[desc.kernels[k].src].astype(np.float64)
    if khat is None:
        khat = np.fft.rfft2(kernel_torus64(desc.kernels[k], w, h))
    return np.fft.irfft2(np.fft.rfft2(src) * khat, s=(h, w))


def kernel_hats64(desc: LeniaDesc) -> list[np.ndarray]:
    return [np.fft.rfft2(kernel_torus64(desc.kernels[k], desc.width, desc.height))
            for k in range(desc.n_kernels)]


def update64(desc: LeniaDesc, u: np.ndarray, s: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(s, np.float32)
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty_like(s)
    lib().or_update64(ctypes.byref(desc), _dp(u), _fp(s), _fp(out))
    return out


def step64(desc: LeniaDesc, s: np.ndarray, method: str = "fft", khats=None) -> np.ndarray:
    """One FP64 oracle step: FP64 convolution (direct or FFT), FP64 growth/update."""
    s = np.ascontiguousarray(s, np.float32).reshape(desc.channels, desc.height, desc.width)
    if method == "direct":
        out = np.empty_like(s)
        lib().or_step_direct64(ctypes.byref(desc), _fp(s), _fp(out))
        return out
    if khats is None:
        khats = kernel_hats64(desc)
    u = np.stack([conv_fft64(desc, s, k, khats[k]) for k in range(desc.n_kernels)])
    return update64(desc, u, s)


def potentials64(desc: LeniaDesc, s: np.ndarray, khats=None) -> np.ndarray:
    if khats is None:
        khats = kernel_hats64(desc)
    s = s.reshape(desc.channels, desc.height, desc.width)
    return np.stack([conv_fft64(desc, s, k, khats[k]) for k in range(desc.n_kernels)])


def stats(s: np.ndarray, channels: int) -> Stats:
    s = np.ascontiguousarray(s, np.float32)
    st = Stats()
    lib().or_stats(_fp(s), s.size // channels, channels, ctypes.byref(st))
    return st


def set_threads(n: int) -> None:
    lib().or_set_threads(n)


def max_threads() -> int:
    return lib().or_max_threads()
