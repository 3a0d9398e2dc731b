"""Rule descriptors: the reference Lenia rule and the benchmark workloads.

`lenia_reference` is the rule the reference registers as "lenia"
(crates/auto-core/src/automata/spectral.rs:286-352,482-490: R=13, one ring,
mu=0.15, sigma=0.015, dt=0.1, Orbium seed), with sizes rounded up to powers
of two as automata::fit_size does (crates/auto-core/src/automata/mod.rs:51-61).
The other builders are the study workloads:

  single_ring_256   256^2, R=13, centred 64^2 random patch, 100 steps
  rings3_4096       4096^2, R=26, rings beta=[1,2/3,1/3], mu=.26, sigma=.036,
                    dt=.1, 64 random 48^2 blobs, 1000 steps
  multichannel_1024 1024^2 x 3 channels, 9 kernels drawn from SplitMix64
  uneven_worlds     384 single-channel worlds of sides 128..1024
Parameter draws use SplitMix64 (crates/auto-life-gpu/src/world.rs:153-159)
with u01(x) = (x >> 11) * 2^-53.
"""
from __future__ import annotations

from ._abi import (LeniaDesc, INIT_ORBIUM, INIT_PATCH, INIT_BLOBS, INIT_DISCS, INIT_NOISE, PATHS, MAX_KERNELS,
                   SHAPE_SMOOTH_DISC, SHAPE_SMOOTH_RING, SHAPE_BAND, RULES)

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream (crates/auto-life-gpu/src/world.rs:153-159)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def u01(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, n: int) -> int:
        return self.next() % n


def _pow2(n: int) -> int:
    return 1 << max(0, (int(n) - 1).bit_length())


def desc(width: int, height: int, kernels: list[dict], dt: float, path: str = "fft_fp32",
         channels: int = 1, init: int = INIT_PATCH, init_size: int = 0, init_count: int = 0,
         rule: str = "lenia") -> LeniaDesc:
    d = LeniaDesc()
    d.width, d.height, d.channels = width, height, channels
    if not 1 <= len(kernels) <= MAX_KERNELS:
        raise ValueError("1..9 kernels")
    d.n_kernels = len(kernels)
    for i, k in enumerate(kernels):
        kd = d.kernels[i]
        kd.radius = k["radius"]
        beta = list(k.get("beta", [1.0] if k.get("shape", 0) == 0 else []))
        kd.rings = len(beta)
        for j, b in enumerate(beta):
            kd.beta[j] = b
        kd.mu, kd.sigma = k.get("mu", 0.0), k.get("sigma", 0.0)
        kd.weight = k.get("weight", 1.0)
        kd.src, kd.dst = k.get("src", 0), k.get("dst", 0)
        kd.shape, kd.inner, kd.aa = k.get("shape", 0), k.get("inner", 0.0), k.get("aa", 0.0)
    d.dt = dt
    d.path = PATHS[path] if isinstance(path, str) else int(path)
    d.init, d.init_size, d.init_count = init, init_size, init_count
    d.rule = RULES[rule] if isinstance(rule, str) else int(rule)
    return d


def lenia_reference(width: int, height: int, path: str = "fft_fp32") -> LeniaDesc:
    """The reference "lenia" entry (spectral.rs:482-490 -> lenia(w, h, 13.0))."""
    return desc(_pow2(width), _pow2(height), [dict(radius=13.0, mu=0.15, sigma=0.015)], 0.1,
                path, init=INIT_ORBIUM)


def smoothlife(width: int, height: int, ra: float = 21.0) -> LeniaDesc:
    """The reference "smoothlife" entry (spectral.rs:354-413, built with
    ra = 21 at :499): inner disc of radius ra/3 and the ring out to ra, both
    antialiased over 0.5 cells, dt = 0.1, scattered-disc seed."""
    ri, aa = ra / 3.0, 0.5
    return desc(_pow2(width), _pow2(height),
                [dict(shape=SHAPE_SMOOTH_DISC, radius=ri, aa=aa),
                 dict(shape=SHAPE_SMOOTH_RING, radius=ra, inner=ri, aa=aa)], 0.1,
                init=INIT_DISCS, rule="smoothlife")


def mnca(width: int, height: int, rings=(12.0, 18.0, 3.0, 8.0)) -> LeniaDesc:
    """The reference "mnca" entry (spectral.rs:415-480; rings [12,18) and
    [3,8) by default, :427), noise seed."""
    i0, o0, i1, o1 = (float(v) for v in rings)
    return desc(_pow2(width), _pow2(height),
                [dict(shape=SHAPE_BAND, radius=o0, inner=i0), dict(shape=SHAPE_BAND, radius=o1, inner=i1)], 1.0,
                init=INIT_NOISE, rule="mnca")


def single_ring_256(path: str = "fft_fp32") -> LeniaDesc:
    """256^2, R=13, mu=.15, sigma=.015, dt=.1, centred 64^2 patch."""
    return desc(256, 256, [dict(radius=13.0, mu=0.15, sigma=0.015)], 0.1, path,
                init=INIT_PATCH, init_size=64)


RINGS3 = dict(radius=26.0, beta=[1.0, 2.0 / 3.0, 1.0 / 3.0], mu=0.26, sigma=0.036)


def rings3_4096(path: str = "fft_fp32", side: int = 4096) -> LeniaDesc:
    return desc(side, side, [dict(RINGS3)], 0.1, path, init=INIT_BLOBS, init_size=48, init_count=64)


def multichannel_kernels(seed: int) -> list[dict]:
    """Nine kernels, k = 3*src + dst, drawn in this order from one SplitMix64 stream."""
    rng = SplitMix64(seed)
    ks = []
    for k in range(9):
        radius = 8 + rng.below(13)
        rings = 1 + rng.below(3)
        beta = [0.2 + 0.8 * rng.u01() for _ in range(rings)]
        mu = 0.1 + 0.25 * rng.u01()
        sigma = 0.01 + 0.04 * rng.u01()
        weight = 0.2 + 0.8 * rng.u01()
        ks.append(dict(radius=float(radius), beta=beta, mu=mu, sigma=sigma, weight=weight,
                       src=k // 3, dst=k % 3))
    return ks


def multichannel_1024(seed: int = 1, path: str = "fft_fp32", side: int = 1024) -> LeniaDesc:
    return desc(side, side, multichannel_kernels(seed), 0.2, path, channels=3,
                init=INIT_PATCH, init_size=side // 4)


def uneven_worlds(seed: int = 1, n: int = 384, path: str = "fft_fp32",
                  sides=(128, 256, 512, 1024)) -> list[tuple[LeniaDesc, int]]:
    """Per world: its own SplitMix64 seed, side, radius, rings, mu, sigma, dt and patch."""
    top = SplitMix64(seed)
    worlds = []
    for _ in range(n):
        wseed = top.next() & 0xFFFFFFFF
        rng = SplitMix64(wseed)
        side = sides[rng.below(len(sides))]
        radius = 8 + rng.below(17)
        rings = 1 + rng.below(3)
        beta = [0.2 + 0.8 * rng.u01() for _ in range(rings)]
        mu = 0.1 + 0.25 * rng.u01()
        sigma = 0.01 + 0.04 * rng.u01()
        dt = (0.05, 0.1, 0.2)[rng.below(3)]
        d = desc(side, side, [dict(radius=float(radius), beta=beta, mu=mu, sigma=sigma)], dt, path,
                 init=INIT_PATCH, init_size=side // 4)
        worlds.append((d, wseed))
    return worlds
