"""Distinct sims on distinct host threads (the C ABI's threading contract:
calls are not re-entrant per handle, but separate handles may run
concurrently).  ctypes releases the GIL around each engine call, so the
threads' creations, autotuning, steps and reads overlap on the device.
Each thread's result must equal the same sim run alone, bit for bit."""
import threading

import numpy as np
import pytest

from auto_amd import Sim, workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    ("fft", lambda: W.rings3_4096(side=1024), 3),
    ("direct", lambda: W.rings3_4096(side=512, path="direct_bf16"), 5),
    ("multichannel", lambda: W.multichannel_1024(seed=2, side=256, path="direct_bf16"), 7),
    ("smoothlife", lambda: W.smoothlife(512, 512), 9),
]


def _run(make, seed, steps=20):
    """State after `steps` steps.  A direct sim may pick another kernel
    variant when its autotuning overlaps other threads' work; every variant
    gives the same bits."""
    sim = Sim(make(), seed)
    sim.step(steps)
    return sim.read_field("state")


def test_sims_on_threads_match_sequential_runs():
    alone = {name: _run(make, seed) for name, make, seed in CASES}
    got, errors = {}, []

    def worker(name, make, seed):
        try:
            for _ in range(2):
                got.setdefault(name, []).append(_run(make, seed))
        except Exception as e:  # reported below; a thread's exception is otherwise lost
            errors.append(f"{name}: {e!r}")

    threads = [threading.Thread(target=worker, args=c) for c in CASES]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not [t for t in threads if t.is_alive()], "a worker thread did not finish"
    assert not errors, errors
    for name, state in alone.items():
        assert len(got[name]) == 2, name
        for i, s in enumerate(got[name]):
            assert np.array_equal(s, state), (name, i)
