"""FP64 oracle of the learned k-step surrogate.  TEST INFRASTRUCTURE.

Only tests/ import this module, to check the GPU network's forward pass and
gradients.  The reference has no learned part (it is new in this engine),
so the network is defined here and in auto_amd/csrc/surrogate.hip
from the same written specification:

  x      [N][H][W]   float32 state (torus), H and W even
  z1 = conv5(x; W1 1->32) + b1,     a1 = gelu(z1)          full resolution
  p  = avgpool2(a1)                                         half resolution
  z2 = conv5(p; W2 32->64) + b2,    a2 = gelu(z2)          half resolution
  u  = upsample2(a2)  (nearest)                             full resolution
  z3 = conv5(u; W3 64->32) + b3,    a3 = gelu(z3)
  s  = a3 + a1                      (skip connection)
  r  = conv5(s; W4 32->1) + b4
  y  = clip(x + r, 0, 1)            (residual output)

conv5 is a 5x5 cross-correlation with circular padding:
  conv5(X; W)[co][y][x] = sum_{ci, dy, dx in -2..2} W[co][(dy+2)*5+(dx+2)][ci] X[ci][y+dy][x+dx]
gelu(z) = z * Phi(z) (the erf form).  Loss = mean((y - target)^2).
The clip's derivative is 1 where 0 < x + r < 1 and 0 elsewhere.

Parameters are one flat float32 vector in this order: W1 [32][25][1], b1 [32],
W2 [64][25][32], b2 [64], W3 [32][25][64], b3 [32], W4 [1][25][32], b4 [1]
(104,129 values).

With bf16_points=True the oracle rounds to bfloat16 where the GPU stores
bf16: the weights, the input x of the first convolution, the activations a1,
p, a2 and s (s = bf16(gelu(z3) + a1)), the gelu'(z) values the forward pass
keeps for the backward pass, and the gradients dz3, dz2 that feed MFMA
convolutions.  Everything else (pre-activations z, r, the backward
input gradients ds, du, dp, dz1) stays in FP32 on the GPU and FP64 here.  The analytic backward is pinned by central finite differences of the
FP64 loss (tests/test_surrogate_oracle.py).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

C1, C2 = 32, 64
SHAPES = [("W1", (C1, 25, 1)), ("b1", (C1,)), ("W2", (C2, 25, C1)), ("b2", (C2,)),
          ("W3", (C1, 25, C2)), ("b3", (C1,)), ("W4", (1, 25, C1)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest even) via float32, as the GPU does."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """x[..., y+dy, x+dx] on the torus (last two axes)."""
    return np.roll(np.roll(x, -dy, axis=-2), -dx, axis=-1)


def conv5(X: np.ndarray, W: np.ndarray) -> np.ndarray:
    """X [N][Cin][H][W], W [Cout][25][Cin] -> [N][Cout][H][W]."""
    out = np.zeros((X.shape[0], W.shape[0]) + X.shape[2:])
    for t in range(25):
        dy, dx = t // 5 - 2, t % 5 - 2
        out += np.einsum("oc,nchw->nohw", W[:, t, :], _shift(X, dy, dx))
    return out


def conv5_input_grad(G: np.ndarray, W: np.ndarray) -> np.ndarray:
    """d/dX of sum(G * conv5(X, W)): G [N][Cout][H][W] -> [N][Cin][H][W]."""
    out = np.zeros((G.shape[0], W.shape[2]) + G.shape[2:])
    for t in range(25):
        dy, dx = t // 5 - 2, t % 5 - 2
        out += _shift(np.einsum("oc,nohw->nchw", W[:, t, :], G), -dy, -dx)
    return out


def conv5_weight_grad(G: np.ndarray, X: np.ndarray) -> np.ndarray:
    """d/dW of sum(G * conv5(X, W)) -> [Cout][25][Cin]."""
    out = np.zeros((G.shape[1], 25, X.shape[1]))
    for t in range(25):
        dy, dx = t // 5 - 2, t % 5 - 2
        out[:, t, :] = np.einsum("nohw,nchw->oc", G, _shift(X, dy, dx))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def pool2(a):
    n, c, h, w = a.shape
    return a.reshape(n, c, h // 2, 2, w // 2, 2).mean(axis=(3, 5))


def up2(a):
    return a.repeat(2, axis=-2).repeat(2, axis=-1)


def forward(theta: np.ndarray, x: np.ndarray, bf16_points: bool = False) -> tuple[np.ndarray, dict]:
    """x [N][H][W] -> (y [N][H][W], cache)."""
    P = unpack(theta)
    rd = bf16 if bf16_points else (lambda v: np.asarray(v, np.float64))
    W = {k: rd(P[k]) if k.startswith("W") else P[k] for k in P}
    x = np.asarray(x, np.float64)
    x_in = rd(x)[:, None]
    z1 = conv5(x_in, W["W1"]) + W["b1"][None, :, None, None]
    a1 = rd(gelu(z1))
    p = rd(pool2(a1))
    z2 = conv5(p, W["W2"]) + W["b2"][None, :, None, None]
    a2 = rd(gelu(z2))
    u = up2(a2)
    z3 = conv5(u, W["W3"]) + W["b3"][None, :, None, None]
    s = rd(gelu(z3) + a1)
    r = (conv5(s, W["W4"]) + W["b4"][None, :, None, None])[:, 0]
    pre = x + r
    y = np.clip(pre, 0.0, 1.0)
    cache = dict(W=W, x=x, x_in=x_in, z1=z1, a1=a1, p=p, z2=z2, a2=a2, u=u, z3=z3, s=s, r=r, pre=pre)
    return y, cache


def loss_and_grad(theta: np.ndarray, x: np.ndarray, target: np.ndarray, bf16_points: bool = False):
    """MSE loss and its gradient with respect to the flat parameter vector."""
    y, c = forward(theta, x, bf16_points)
    rd = bf16 if bf16_points else (lambda v: v)
    diff = y - np.asarray(target, np.float64)
    loss = float(np.mean(diff ** 2))
    dy = 2.0 * diff / diff.size
    dr = dy * ((c["pre"] > 0.0) & (c["pre"] < 1.0))
    W = c["W"]
    g = {}
    dr4 = dr[:, None]
    g["b4"] = dr.sum(axis=(0, 1, 2), keepdims=False).reshape(1)
    g["W4"] = conv5_weight_grad(dr4, c["s"])
    ds = conv5_input_grad(dr4, W["W4"])
    dz3 = rd(ds * rd(gelu_grad(c["z3"])))
    g["b3"] = dz3.sum(axis=(0, 2, 3))
    g["W3"] = conv5_weight_grad(dz3, c["u"])
    du = conv5_input_grad(dz3, W["W3"])
    n, ch, h, w = du.shape
    da2 = du.reshape(n, ch, h // 2, 2, w // 2, 2).sum(axis=(3, 5))
    dz2 = rd(da2 * rd(gelu_grad(c["z2"])))
    g["b2"] = dz2.sum(axis=(0, 2, 3))
    g["W2"] = conv5_weight_grad(dz2, c["p"])
    dp = conv5_input_grad(dz2, W["W2"])
    dz1 = (ds + up2(dp) / 4.0) * rd(gelu_grad(c["z1"]))
    g["b1"] = dz1.sum(axis=(0, 2, 3))
    g["W1"] = conv5_weight_grad(dz1, c["x_in"])
    return loss, pack(g)


def init_params(seed: int) -> np.ndarray:
    """The GPU's initialisation, restated: SplitMix64 uniforms scaled to
    +-sqrt(3/fan_in) (unit-variance inputs keep unit variance), biases 0 and
    the last layer scaled by 0.1 so the untrained net starts near identity."""
    from auto_amd.workloads import SplitMix64
    rng = SplitMix64(seed)
    out = []
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        if name.startswith("b"):
            out.append(np.zeros(n))
            continue
        fan_in = shape[1] * shape[2]
        lim = np.sqrt(3.0 / fan_in) * (0.1 if name == "W4" else 1.0)
        out.append(np.array([(2.0 * rng.u01() - 1.0) * lim for _ in range(n)]))
    return np.concatenate(out).astype(np.float32).astype(np.float64)
