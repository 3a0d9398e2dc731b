"""ctypes mirror of include/auto_hip.h and the loader for libauto_hip.so.

The shared library is the product; this module only describes its C ABI.
Loading fails loudly when the library is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

MAX_CHANNELS = 3
MAX_KERNELS = 9
MAX_RINGS = 3

OK = 0
ERR_INVALID = 1
ERR_DEVICE = 2
ERR_ALLOC = 3
ERR_UNKNOWN_FIELD = 4
ERR_SIZE = 5
ERR_UNSUPPORTED = 6
ERR_DIVERGED = 7

PATH_FFT_FP32 = 0
PATH_DIRECT_BF16 = 1
PATH_DIRECT_BF16_X1 = 2
PATH_DIRECT_BF16_X3 = 3
PATH_AUTO = 4
BATCH_FFT_ABOVE = 30

INIT_ORBIUM = 0
INIT_PATCH = 1
INIT_BLOBS = 2
INIT_ZERO = 3
INIT_DISCS = 4
INIT_NOISE = 5

SHAPE_LENIA = 0
SHAPE_SMOOTH_DISC = 1
SHAPE_SMOOTH_RING = 2
SHAPE_BAND = 3

RULE_LENIA = 0
RULE_SMOOTHLIFE = 1
RULE_MNCA = 2
RULES = {"lenia": RULE_LENIA, "smoothlife": RULE_SMOOTHLIFE, "mnca": RULE_MNCA}

PATHS = {"fft_fp32": PATH_FFT_FP32, "direct_bf16": PATH_DIRECT_BF16,
         "direct_bf16_x1": PATH_DIRECT_BF16_X1, "direct_bf16_x3": PATH_DIRECT_BF16_X3, "auto": PATH_AUTO}


class Kernel(ctypes.Structure):
    _fields_ = [
        ("radius", ctypes.c_float),
        ("rings", ctypes.c_uint32),
        ("beta", ctypes.c_float * MAX_RINGS),
        ("mu", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("weight", ctypes.c_float),
        ("src", ctypes.c_uint32),
        ("dst", ctypes.c_uint32),
        ("shape", ctypes.c_uint32),
        ("inner", ctypes.c_float),
        ("aa", ctypes.c_float),
    ]


class LeniaDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("channels", ctypes.c_uint32),
        ("n_kernels", ctypes.c_uint32),
        ("kernels", Kernel * MAX_KERNELS),
        ("dt", ctypes.c_float),
        ("path", ctypes.c_uint32),
        ("init", ctypes.c_uint32),
        ("init_size", ctypes.c_uint32),
        ("init_count", ctypes.c_uint32),
        ("rule", ctypes.c_uint32),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("channels", ctypes.c_uint32),
        ("n_kernels", ctypes.c_uint32),
        ("path", ctypes.c_uint32),
        ("cells_per_step", ctypes.c_uint64),
        ("generations_per_step", ctypes.c_uint64),
        ("total_bytes", ctypes.c_uint64),
        ("generation", ctypes.c_uint64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("mass", ctypes.c_double * MAX_CHANNELS),
        ("min", ctypes.c_float * MAX_CHANNELS),
        ("max", ctypes.c_float * MAX_CHANNELS),
        ("nonfinite", ctypes.c_uint64),
        ("generation", ctypes.c_uint64),
    ]


class DirectTile(ctypes.Structure):
    _fields_ = [("cols", ctypes.c_uint32), ("rows", ctypes.c_uint32), ("variant", ctypes.c_uint32)]


class KernelTime(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 40),
        ("launches", ctypes.c_uint64),
        ("total_ms", ctypes.c_double),
    ]


BATCH_MAX_CANDIDATES = 8
BATCH_MAX_GROUP_TUNE = 64


class BatchTune(ctypes.Structure):
    _fields_ = [("tile_cols", ctypes.c_uint32), ("tile_rows", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("ms_per_step", ctypes.c_double)]


class BatchGroupTune(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("column", ctypes.c_uint32),
                ("row", ctypes.c_uint32), ("ms_per_step", ctypes.c_double)]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("n_worlds", ctypes.c_uint32),
        ("tile_cols", ctypes.c_uint32),
        ("tile_rows", ctypes.c_uint32),
        ("grid", ctypes.c_uint32),
        ("ntiles", ctypes.c_uint32),
        ("graph_steps", ctypes.c_uint32),
        ("total_cells", ctypes.c_uint64),
        ("total_bytes", ctypes.c_uint64),
        ("generation", ctypes.c_uint64),
        ("n_tune", ctypes.c_uint32),
        ("tune", BatchTune * BATCH_MAX_CANDIDATES),
        ("fft_worlds", ctypes.c_uint32),
        ("fft_groups", ctypes.c_uint32),
        ("n_group_tune", ctypes.c_uint32),
        ("group_tune", BatchGroupTune * BATCH_MAX_GROUP_TUNE),
    ]


SURROGATE_PARAMS = 104129


class SurrogateReport(ctypes.Structure):
    _fields_ = [
        ("lenia_steps", ctypes.c_uint64),
        ("surrogate_calls", ctypes.c_uint64),
        ("plain_steps", ctypes.c_uint64),
        ("judges", ctypes.c_uint64),
        ("fallbacks", ctypes.c_uint64),
        ("max_judge_rel_l2", ctypes.c_double),
        ("last_judge_rel_l2", ctypes.c_double),
        ("ms", ctypes.c_double),
    ]


LIB_NAME = "libauto_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# Every symbol include/auto_hip.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_SIGNATURES = {
    "ah_lenia_build": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(LeniaDesc)]),
    "ah_sim_new": (ctypes.c_int, [ctypes.POINTER(LeniaDesc), ctypes.c_uint32, ctypes.POINTER(_P)]),
    "ah_sim_step": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "ah_sim_wait": (ctypes.c_int, [_P]),
    "ah_sim_read_field": (ctypes.c_int, [_P, ctypes.c_char_p, _P, ctypes.c_size_t]),
    "ah_sim_write_field": (ctypes.c_int, [_P, ctypes.c_char_p, _P, ctypes.c_size_t]),
    "ah_sim_stats": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "ah_sim_check": (ctypes.c_int, [_P]),
    "ah_state_stats_noalloc": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint32, _P, ctypes.POINTER(Stats)]),
    "ah_sim_info": (ctypes.c_int, [_P, ctypes.POINTER(Info)]),
    "ah_sim_destroy": (None, [_P]),
    "ah_last_error": (ctypes.c_char_p, []),
    "ah_sim_profile": (ctypes.c_int, [_P, ctypes.c_int]),
    "ah_sim_direct_tile": (ctypes.c_int, [_P, ctypes.POINTER(DirectTile)]),
    "ah_sim_set_direct_tile": (ctypes.c_int, [_P, ctypes.POINTER(DirectTile)]),
    "ah_sim_profile_read": (ctypes.c_int, [_P, ctypes.POINTER(KernelTime), ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint32)]),
    "ah_copy_probe": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "ah_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "ah_sim_render": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t]),
    "ah_batch_new": (ctypes.c_int, [ctypes.POINTER(LeniaDesc), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.POINTER(_P)]),
    "ah_batch_step": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "ah_batch_wait": (ctypes.c_int, [_P]),
    "ah_batch_read_world": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_size_t]),
    "ah_batch_write_world": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_size_t]),
    "ah_batch_info_get": (ctypes.c_int, [_P, ctypes.POINTER(BatchInfo)]),
    "ah_batch_measure": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "ah_batch_destroy": (None, [_P]),
    "ah_surrogate_new": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(_P)]),
    "ah_surrogate_destroy": (None, [_P]),
    "ah_surrogate_get_params": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "ah_surrogate_set_params": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "ah_surrogate_forward": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P]),
    "ah_surrogate_loss_grad": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.POINTER(ctypes.c_double), _P]),
    "ah_surrogate_collect": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32]),
    "ah_surrogate_dataset": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, _P, _P]),
    "ah_surrogate_train": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                          ctypes.c_uint32, _P]),
    "ah_surrogate_eval": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "ah_surrogate_advance": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                            ctypes.POINTER(SurrogateReport)]),
}

_lib = None


class EngineError(RuntimeError):
    """A non-OK ah_status, carrying the code and ah_last_error()'s message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"auto_hip error {code}: {message}")
        self.code = code


def load() -> ctypes.CDLL:
    """Load libauto_hip.so from the package directory (or the path in the
    AUTO_HIP_LIB environment variable); raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("AUTO_HIP_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != OK:
        msg = load().ah_last_error()
        raise EngineError(code, msg.decode() if msg else "")


def symbols() -> list[str]:
    return list(_SIGNATURES)
