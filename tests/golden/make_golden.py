"""Writes the small golden fixtures under tests/golden/.

The fixtures are inputs and FP64 outputs of the oracle at small sizes:
  kat_lenia_128x64.npz  the reference's known-answer case (automata.rs:300-362):
                        Orbium seed, one step, FP64 direct and FP32 reference step
  rings3_128.npz        three-ring shell (R=20), 64 blobs, one FP64 direct step
  multichannel_64.npz   3 channels / 9 kernels, one FP64 direct step
  dft_32.npz            complex rows and their brute-force FP64 DFT
The FP64 direct values are cross-checked by tests/test_oracle.py against an
independent FP64 FFT convolution before they are trusted.
Run from the repo root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from auto_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    d = W.lenia_reference(128, 64)
    seed = next(s for s in range(1, 64) if (O.init_state(d, s) > 0).any())
    s0 = O.init_state(d, seed)
    np.savez_compressed(os.path.join(OUT, "kat_lenia_128x64.npz"), seed=seed, s0=s0[0],
                        step64=O.step64(d, s0, "direct")[0], step_ref32=O.step_ref32(d, s0)[0])

    d = W.rings3_4096(side=128)
    d.kernels[0].radius = 20.0
    s0 = O.init_state(d, 17)
    np.savez_compressed(os.path.join(OUT, "rings3_128.npz"), seed=17, s0=s0, step64=O.step64(d, s0, "direct"))

    d = W.multichannel_1024(seed=2, side=64)
    s0 = O.init_state(d, 4)
    np.savez_compressed(os.path.join(OUT, "multichannel_64.npz"), rule_seed=2, seed=4, s0=s0,
                        step64=O.step64(d, s0, "direct"))

    rng = np.random.default_rng(32)
    x = (rng.standard_normal((4, 32)) + 1j * rng.standard_normal((4, 32))).astype(np.complex64)
    j = np.arange(32)
    X = x.astype(np.complex128) @ np.exp(-2j * np.pi * np.outer(j, j) / 32).T
    np.savez_compressed(os.path.join(OUT, "dft_32.npz"), x=x, X=X)


if __name__ == "__main__":
    main()
