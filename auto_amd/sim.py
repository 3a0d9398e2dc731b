"""Sim: the reference's engine interface over libauto_hip.so.

Mirrors crates/auto-core/src/sim.rs: `Sim::new(gpu, program, seed)` (:202),
`step(gpu, steps) -> generations` (:297-309, asynchronous), `read_field` /
`write_field` (:559-595) with row-major little-endian f32 bytes, and the
Program metadata `cells_per_step`, `generations_per_step`, `total_bytes()`
(crates/auto-core/src/program.rs:101-129).  Errors raise EngineError with the
engine's message, where the reference returns Err(String) or panics.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import LeniaDesc, Info, Stats, KernelTime, DirectTile, check, load
from . import workloads


def available() -> bool:
    """True when the built library loads (it does not imply a GPU is present)."""
    try:
        load()
        return True
    except (ImportError, OSError):
        return False


def build(id: str, width: int, height: int) -> LeniaDesc:
    """automata::build (crates/auto-core/src/automata/mod.rs:63-95) through the C registry."""
    d = LeniaDesc()
    check(load().ah_lenia_build(id.encode(), width, height, ctypes.byref(d)))
    return d


class Sim:
    def __init__(self, desc: LeniaDesc, seed: int = 1):
        lib = load()
        self._lib = lib
        self.desc = desc
        h = ctypes.c_void_p()
        check(lib.ah_sim_new(ctypes.byref(desc), seed & 0xFFFFFFFF, ctypes.byref(h)))
        self._h = h
        self.seed = seed
        info = self.info()
        self.width, self.height, self.channels = info.width, info.height, info.channels
        self.cells_per_step = info.cells_per_step
        self.generations_per_step = info.generations_per_step

    # -- lifecycle --------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ah_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping ---------------------------------------------------------
    def step(self, steps: int = 1) -> int:
        gens = ctypes.c_uint64()
        check(self._lib.ah_sim_step(self._h, steps, ctypes.byref(gens)))
        return gens.value

    def wait(self) -> None:
        check(self._lib.ah_sim_wait(self._h))

    # -- fields -----------------------------------------------------------
    def field_shape(self, name: str) -> tuple:
        if name == "state":
            return (self.channels, self.height, self.width)
        return (self.height, self.width)

    def read_field(self, name: str = "state") -> np.ndarray:
        out = np.empty(self.field_shape(name), np.float32)
        check(self._lib.ah_sim_read_field(self._h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def write_field(self, name: str, data: np.ndarray) -> None:
        a = np.ascontiguousarray(data, np.float32)
        check(self._lib.ah_sim_write_field(self._h, name.encode(), a.ctypes.data, a.nbytes))

    # -- metadata ---------------------------------------------------------
    def info(self) -> Info:
        i = Info()
        check(self._lib.ah_sim_info(self._h, ctypes.byref(i)))
        return i

    def total_bytes(self) -> int:
        return self.info().total_bytes

    def stats(self) -> Stats:
        s = Stats()
        check(self._lib.ah_sim_stats(self._h, ctypes.byref(s)))
        return s

    def check(self) -> None:
        """Raise EngineError (AH_ERR_DIVERGED) if any cell is NaN or infinite."""
        check(self._lib.ah_sim_check(self._h))

    # -- view -------------------------------------------------------------
    def render(self, out_w: int, out_h: int, center=None, scale: float | None = None) -> np.ndarray:
        """rgba8 [out_h][out_w][4] view of channel 0 (Sim::render, sim.rs:452-514).
        Without a camera the world is fitted to the frame (Camera::fit,
        sim.rs:52-61)."""
        if scale is None:
            scale = max(self.width / out_w, self.height / out_h)
        if center is None:
            center = (self.width / 2.0, self.height / 2.0)
        out = np.empty((out_h, out_w, 4), np.uint8)
        check(self._lib.ah_sim_render(self._h, out_w, out_h, float(center[0]), float(center[1]), float(scale),
                                      out.ctypes.data, out.nbytes))
        return out

    # -- direct-path kernel variant ----------------------------------------
    def direct_tile(self) -> tuple:
        """(cols, rows, variant) of the direct-path kernel in use: variant 0
        tile kernel, 1 ring kernel, 2 one workgroup per (tile, kernel)."""
        t = DirectTile()
        check(self._lib.ah_sim_direct_tile(self._h, ctypes.byref(t)))
        return (t.cols, t.rows, t.variant)

    def set_direct_tile(self, cols: int, rows: int, variant: int = 0) -> None:
        """Force a direct-path variant (raises if not admissible here)."""
        t = DirectTile(cols, rows, variant)
        check(self._lib.ah_sim_set_direct_tile(self._h, ctypes.byref(t)))

    # -- kernel timing ----------------------------------------------------
    def profile(self, enable: bool = True, every: int = 1) -> None:
        """Time kernel launches with HIP events: every launch, or every
        `every`-th launch of each kernel (keeps the events' cost out of the
        step rate).  Enabling resets the counters."""
        check(self._lib.ah_sim_profile(self._h, max(1, int(every)) if enable else 0))

    def profile_read(self) -> dict:
        cap = 32
        arr = (KernelTime * cap)()
        n = ctypes.c_uint32()
        check(self._lib.ah_sim_profile_read(self._h, arr, cap, ctypes.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}


class Batch:
    """Hundreds of uneven single-channel worlds advanced together (ah_batch_*).

    `worlds` is a list of (LeniaDesc, seed); every desc uses path
    "direct_bf16".  No counterpart in the reference, which runs one world per
    Sim; `read_world(i)` returns what `Sim.read_field("state")[0]` would.
    """

    def __init__(self, worlds):
        lib = load()
        self._lib = lib
        n = len(worlds)
        descs = (LeniaDesc * n)(*[d for d, _ in worlds])
        seeds = (ctypes.c_uint32 * n)(*[s & 0xFFFFFFFF for _, s in worlds])
        h = ctypes.c_void_p()
        check(lib.ah_batch_new(descs, n, seeds, ctypes.byref(h)))
        self._h = h
        self.descs = [d for d, _ in worlds]
        self.seeds = [s for _, s in worlds]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ah_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, steps: int = 1) -> None:
        check(self._lib.ah_batch_step(self._h, steps))

    def wait(self) -> None:
        check(self._lib.ah_batch_wait(self._h))

    def read_world(self, i: int) -> np.ndarray:
        d = self.descs[i]
        out = np.empty((d.height, d.width), np.float32)
        check(self._lib.ah_batch_read_world(self._h, i, out.ctypes.data, out.nbytes))
        return out

    def write_world(self, i: int, data: np.ndarray) -> None:
        a = np.ascontiguousarray(data, np.float32)
        check(self._lib.ah_batch_write_world(self._h, i, a.ctypes.data, a.nbytes))

    def info(self) -> _abi.BatchInfo:
        i = _abi.BatchInfo()
        check(self._lib.ah_batch_info_get(self._h, ctypes.byref(i)))
        return i

    def measure(self, steps: int) -> dict:
        k, g, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self._lib.ah_batch_measure(self._h, steps, ctypes.byref(k), ctypes.byref(g), ctypes.byref(e)))
        return {"kernel_ms": k.value, "graph_wall_ms": g.value, "eager_wall_ms": e.value}


class Surrogate:
    """The learned k-step surrogate (ah_surrogate_*); no reference counterpart."""

    def __init__(self, seed: int = 1):
        lib = load()
        self._lib = lib
        h = ctypes.c_void_p()
        check(lib.ah_surrogate_new(seed & 0xFFFFFFFF, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ah_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(_abi.SURROGATE_PARAMS, np.float32)
        check(self._lib.ah_surrogate_get_params(self._h, out.ctypes.data, out.size))
        return out

    def set_params(self, theta: np.ndarray) -> None:
        a = np.ascontiguousarray(theta, np.float32)
        check(self._lib.ah_surrogate_set_params(self._h, a.ctypes.data, a.size))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        n, h, w = x.shape
        y = np.empty_like(x)
        check(self._lib.ah_surrogate_forward(self._h, x.ctypes.data, n, h, w, y.ctypes.data))
        return y

    def loss_grad(self, x: np.ndarray, target: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        n, h, w = x.shape
        loss = ctypes.c_double()
        g = np.empty(_abi.SURROGATE_PARAMS, np.float32)
        check(self._lib.ah_surrogate_loss_grad(self._h, x.ctypes.data, t.ctypes.data, n, h, w, ctypes.byref(loss),
                                               g.ctypes.data))
        return loss.value, g

    def collect(self, sim: Sim, pairs: int, crop: int = 256, horizon: int = 4, seed: int = 1) -> None:
        check(self._lib.ah_surrogate_collect(self._h, sim._h, pairs, crop, horizon, seed))
        self.crop, self.horizon, self.n_pairs = crop, horizon, pairs

    def dataset(self, first: int, count: int):
        x = np.empty((count, self.crop, self.crop), np.float32)
        y = np.empty_like(x)
        check(self._lib.ah_surrogate_dataset(self._h, first, count, x.ctypes.data, y.ctypes.data))
        return x, y

    def train(self, steps: int, batch: int = 64, n_train: int | None = None, lr: float = 1e-3,
              seed: int = 1) -> np.ndarray:
        losses = np.empty(steps, np.float32)
        n_train = self.n_pairs if n_train is None else n_train
        check(self._lib.ah_surrogate_train(self._h, steps, batch, n_train, lr, seed, losses.ctypes.data))
        return losses

    def eval(self, first: int, count: int) -> float:
        mse = ctypes.c_double()
        check(self._lib.ah_surrogate_eval(self._h, first, count, ctypes.byref(mse)))
        return mse.value

    def advance(self, sim: Sim, steps: int, judge_every: int = 16, tol: float = 1e-2) -> dict:
        r = _abi.SurrogateReport()
        check(self._lib.ah_surrogate_advance(self._h, sim._h, steps, judge_every, tol, ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in r._fields_}


__all__ = ["Sim", "Batch", "Surrogate", "build", "available", "workloads", "_abi"]
