"""FP64 restatement of the MLP surrogate (forward, backward, Adam).

TEST INFRASTRUCTURE ONLY: the checker of the gfx950 MLP kernels (spk_mlp_*), used by
tests/ and nowhere in the product path.  The surrogate has no counterpart in the
reference; this file defines its mathematics.

Network: 5 input features -> 4 hidden layers of 256 SiLU units -> 3 outputs.
Parameters (one flat float32 vector per model, in this order):
    W1 [256][5], b1 [256], W2 [256][256], b2, W3 [256][256], b3, W4 [256][256], b4,
    W5 [3][256], b5 [3]                                           (199683 values)
Loss: L = (1/M) sum_m sum_o w_o (y[m][o] - t[m][o])^2.

mode "exact": every quantity in float64.
mode "bf16": the rounding points of the device kernels, with float64 in between:
  * weights are used as bf16(W) (round to nearest even), biases as float32;
  * an input feature f (float32) enters as hi + lo, hi = bf16(f), lo = bf16(f - hi);
  * hidden activations a = bf16(silu(z)) and slopes s = bf16(silu'(z));
  * output gradient dy = bf16(2 w_o (y - t) / M); hidden gradients dz = bf16(da * s);
  * every sum (over k in a layer, over rows in a weight gradient) is exact here and
    float32 on the device, so the two agree to float32 rounding of those sums plus the
    rare bf16 value that the two round differently.
"""
from __future__ import annotations

import numpy as np

H = 256
DIN = 5
DOUT = 3
SHAPES = [("W1", (H, DIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, H)), ("b4", (H,)), ("W5", (DOUT, H)), ("b5", (DOUT,))]
OFFSETS = {}
_o = 0
for _n, _s in SHAPES:
    OFFSETS[_n] = (_o, _s)
    _o += int(np.prod(_s))
P = _o


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) through float32, returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(flat):
    flat = np.asarray(flat, dtype=np.float64)
    return {n: flat[o:o + int(np.prod(s))].reshape(s) for n, (o, s) in OFFSETS.items()}


def pack(d):
    out = np.zeros(P, np.float64)
    for n, (o, s) in OFFSETS.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[n], np.float64).reshape(-1)
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_slope(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, mode="bf16"):
    """Outputs y [M][3] and the cache the backward pass needs."""
    p = unpack(params)
    if mode == "bf16":
        W = {k: bf16(p[k]) for k in ("W1", "W2", "W3", "W4", "W5")}
        p = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}   # float32 biases
        xf = np.asarray(x, np.float32)
        hi = bf16(xf)
        lo = bf16(xf.astype(np.float64) - hi)       # exact in float32
        a = hi + lo
        rnd = bf16
    else:
        W = {k: p[k] for k in ("W1", "W2", "W3", "W4", "W5")}
        a = np.asarray(x, np.float64)
        rnd = lambda v: v
    acts, slopes = [a], []
    for l in range(1, 5):
        z = acts[-1] @ W[f"W{l}"].T + p[f"b{l}"]
        acts.append(rnd(silu(z)))
        slopes.append(rnd(silu_slope(z)))
    y = acts[-1] @ W["W5"].T + p["b5"]
    return y, dict(acts=acts, slopes=slopes, W=W, rnd=rnd)


def loss_and_grads(params, x, t, wout=(1.0, 1.0, 1.0), mode="bf16"):
    """Loss and its gradient with respect to the flat parameter vector."""
    y, c = forward(params, x, mode)
    M = y.shape[0]
    w = np.asarray(wout, np.float64)
    r = y - np.asarray(t, np.float64)
    loss = float((w * r * r).sum() / M)
    rnd, W, acts, slopes = c["rnd"], c["W"], c["acts"], c["slopes"]
    dy = rnd(2.0 * w * r / M)
    g = {"W5": dy.T @ acts[4], "b5": dy.sum(0)}
    da = dy @ W["W5"]
    for l in range(4, 0, -1):
        dz = rnd(da * slopes[l - 1])
        g[f"W{l}"] = dz.T @ acts[l - 1]
        g[f"b{l}"] = dz.sum(0)
        if l > 1:
            da = dz @ W[f"W{l}"]
    return loss, pack(g), y


def lr_at(t, lr0, lr_min, total):
    """Cosine schedule, step t = 1, 2, ...: lr0 at t = 1, lr_min from t = total + 1 on."""
    u = min(max(t - 1, 0), total) / max(total, 1)
    return lr_min + 0.5 * (lr0 - lr_min) * (1.0 + np.cos(np.pi * u))


class Adam:
    """Adam (Kingma & Ba 2015, Algorithm 1) with bias correction, no weight decay."""

    def __init__(self, params, beta1=0.9, beta2=0.999, eps=1e-8):
        self.p = np.asarray(params, np.float64).copy()
        self.m = np.zeros_like(self.p)
        self.v = np.zeros_like(self.p)
        self.t = 0
        self.b1, self.b2, self.eps = beta1, beta2, eps

    def step(self, g, lr):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.p -= lr * mh / (np.sqrt(vh) + self.eps)
        return self.p
