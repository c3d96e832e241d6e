"""NumPy float64 restatement of the warm-start network (forward, MSE loss,
backward) for checking the GPU kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product library.

The network has no counterpart in the reference; this file restates its
definition: conv3x3 layers (zero padding, cross-correlation) 5 -> 32,
`hidden` x (32 -> 32), 32 -> 2, ReLU after every layer but the last, loss
L = sum (Y - T)^2 / (2 M) over M pixels.  Parameters are flattened per layer
as W[out][tap 0..8][in] (tap t = 3*(dy+1) + (dx+1)) followed by bias[out].

`bf16=True` rounds every GEMM operand to bfloat16 (nearest even) where the
GPU kernels do -- layer inputs and weights in the forward pass, the output
gradient and the weights in the data-gradient pass, the output gradient and
the layer input in the weight-gradient pass -- and keeps float64 elsewhere.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_dims(hidden: int, channels: int = 32):
    dims = [(5, channels)] + [(channels, channels)] * hidden + [(channels, 2)]
    return dims


def unflatten(params: np.ndarray, hidden: int, channels: int = 32):
    """[(W[out][9][in], b[out])] per layer from the flat parameter vector."""
    out, off = [], 0
    for cin, cout in layer_dims(hidden, channels):
        n = cout * 9 * cin
        W = np.asarray(params[off:off + n], np.float64).reshape(cout, 9, cin)
        off += n
        b = np.asarray(params[off:off + cout], np.float64)
        off += cout
        out.append((W, b))
    assert off == len(params)
    return out


def im2col(X: np.ndarray, sgn: int = 1) -> np.ndarray:
    """X [B,H,W,C] -> [B,H,W,9,C]: entry t holds X at (y + sgn*dy, x + sgn*dx),
    zero outside, with (dy, dx) = (t // 3 - 1, t % 3 - 1)."""
    B, H, W, C = X.shape
    P = np.zeros((B, H + 2, W + 2, C), X.dtype)
    P[:, 1:-1, 1:-1, :] = X
    cols = np.empty((B, H, W, 9, C), X.dtype)
    for t in range(9):
        dy, dx = sgn * (t // 3 - 1), sgn * (t % 3 - 1)
        cols[:, :, :, t, :] = P[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]
    return cols


def forward(params, X, hidden, bf16=False, channels=32):
    """Returns the list of layer outputs (post-activation), the last = flow."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    acts = []
    a = np.asarray(X, np.float64)
    layers = unflatten(params, hidden, channels)
    for i, (W, b) in enumerate(layers):
        cols = im2col(r(a.astype(np.float32)) if bf16 else a)
        Wr = r(W.astype(np.float32)) if bf16 else W
        y = np.einsum("bhwtc,ntc->bhwn", cols, Wr) + b
        if i + 1 < len(layers):
            y = np.maximum(y, 0.0)
        if bf16:
            y = y.astype(np.float32).astype(np.float64)   # fp32 storage
        acts.append(y)
        a = y
    return acts


def loss_and_grads(params, X, T, hidden, bf16=False, channels=32):
    """T: planar targets [B, 2, H, W].  Returns (loss, flat gradient)."""
    X = np.asarray(X, np.float64)
    B, H, W, _ = X.shape
    M = B * H * W
    acts = forward(params, X, hidden, bf16, channels)
    Tn = np.moveaxis(np.asarray(T, np.float64).reshape(B, 2, H, W), 1, -1)
    d = acts[-1] - Tn
    loss = float(np.sum(d * d) / (2 * M))
    dY = d / M
    if bf16:
        dY = dY.astype(np.float32).astype(np.float64)
    layers = unflatten(params, hidden, channels)
    grads = [None] * len(layers)
    r = bf16_round
    for i in range(len(layers) - 1, -1, -1):
        Wt, _ = layers[i]
        a_in = acts[i - 1] if i else X
        cols = im2col(r(a_in.astype(np.float32)) if bf16 else a_in)
        dYr = r(dY.astype(np.float32)) if bf16 else dY
        gW = np.einsum("bhwn,bhwtc->ntc", dYr, cols)
        gb = dY.sum(axis=(0, 1, 2))
        grads[i] = (gW, gb)
        if i == 0:
            break
        Wr = r(Wt.astype(np.float32)) if bf16 else Wt
        # dX[m][c] = sum_(t,n) dY[m - s(t)][n] W[n][t][c]
        dcols = im2col(dYr, sgn=-1)
        dX = np.einsum("bhwtn,ntc->bhwc", dcols, Wr)
        dX = dX * (acts[i - 1] > 0)
        if bf16:
            dX = dX.astype(np.float32).astype(np.float64)
        dY = dX
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def init_params(hidden, seed=0, channels=32):
    rng = np.random.default_rng(seed)
    parts = []
    dims = layer_dims(hidden, channels)
    for i, (cin, cout) in enumerate(dims):
        sd = np.sqrt(2.0 / (9 * cin)) * (1.0 if i + 1 < len(dims) else 0.1)
        parts.append(rng.normal(0.0, sd, cout * 9 * cin))
        parts.append(rng.normal(0.0, 0.05, cout))
    return np.concatenate(parts).astype(np.float32)
