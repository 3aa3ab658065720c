"""CPU reference of the learned radiance cache, for tests only (a checker,
never called by the library).

The cache has no counterpart in the reference renderer; this restates the
model defined in voxeltoy_amd/csrc/vt_cache.h:
  - encoding: per raw scalar v, (s, c) = sincos(pi v) (the shared vtmath
    routine, through the C oracle), then 12 bands of (s, c) with the
    double-angle recurrence s' = (2s)c, c' = 1 - (2s)s in float32;
  - MLP 288-64-64-64-64-64-3, ReLU hidden layers, linear output;
  - loss: mean over samples and channels of (y - t)^2 / (sg(y)^2 + 0.01);
  - the FP16-operand / FP32-accumulate inference emulated with exact
    float64 products of float16-rounded operands;
  - the mixed-precision training step (forward as the inference, keeping
    the float16 layer inputs; float32 loss gradient; backward and weight
    gradients with the float32 weights and float64 sums) emulated the same
    way, so the matrix-core training kernels can be held to float32
    accumulation-order error.
"""
from __future__ import annotations

import numpy as np

IN, W, OUT, RAW, BANDS = 288, 64, 3, 12, 12
LAYERS = [(IN, W), (W, W), (W, W), (W, W), (W, W), (W, OUT)]
N_PARAMS = sum(i * o + o for i, o in LAYERS)


def unpack(p):
    """Parameter vector -> [(W (out, in), b (out,)), ...] (views)."""
    out, off = [], 0
    for i, o in LAYERS:
        w = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((w, b))
    return out


def encode(raw, sincos):
    """raw (n, 12) float32 -> (n, 288) float32, feature 24j + 2k + {0, 1}."""
    raw = np.asarray(raw, np.float32)
    n = raw.shape[0]
    out = np.zeros((n, IN), np.float32)
    x = (np.float32(np.pi) * raw).astype(np.float32)
    s, c = sincos(x.reshape(-1))
    s = s.reshape(n, RAW).astype(np.float32)
    c = c.reshape(n, RAW).astype(np.float32)
    two, one = np.float32(2), np.float32(1)
    for k in range(BANDS):
        out[:, 2 * k::24] = s
        out[:, 2 * k + 1::24] = c
        s, c = ((two * s) * c).astype(np.float32), (one - (two * s) * s).astype(np.float32)
    return out


def forward64(p, x):
    """float64 forward: returns y (n, 3) and the layer outputs."""
    layers = unpack(np.asarray(p, np.float64))
    h = np.asarray(x, np.float64)
    hs = []
    for li, (w, b) in enumerate(layers):
        z = h @ w.T + b
        h = z if li == 5 else np.maximum(z, 0.0)
        hs.append(h)
    return h, hs


def loss_and_grad64(p, x, t):
    """Relative-L2 loss and its gradient (float64, stop-gradient on the
    denominator)."""
    p = np.asarray(p, np.float64)
    layers = unpack(p)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    acts = [x]
    zs = []
    h = x
    for li, (w, b) in enumerate(layers):
        z = h @ w.T + b
        zs.append(z)
        h = z if li == 5 else np.maximum(z, 0.0)
        acts.append(h)
    y = h
    wgt = 1.0 / (y * y + 0.01)
    loss = np.sum((y - t) ** 2 * wgt) / (OUT * n)
    dz = 2.0 * (y - t) * wgt / (OUT * n)
    grads = []
    for li in range(5, -1, -1):
        w, _ = layers[li]
        a_in = acts[li]
        gw = dz.T @ a_in
        gb = dz.sum(0)
        grads.append((gw, gb))
        if li > 0:
            dh = dz @ w
            dz = dh * (zs[li - 1] > 0)
    grads.reverse()
    g = np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])
    return loss, g


def forward_fp16(p, x):
    """Emulation of the matrix-core inference: float16 weights and layer
    inputs, exact products, float64 sums, float32 bias add, ReLU, float16
    rounding of every hidden activation."""
    layers = unpack(np.asarray(p, np.float32))
    h = np.asarray(x, np.float32).astype(np.float16).astype(np.float64)
    for li, (w, b) in enumerate(layers):
        w16 = w.astype(np.float16).astype(np.float64)
        z = (h @ w16.T).astype(np.float32) + b.astype(np.float32)
        if li == 5:
            return z.astype(np.float32)
        h = np.maximum(z, np.float32(0)).astype(np.float16).astype(np.float64)
    return h


def _to_fp16(a):
    return np.asarray(a).astype(np.float16).astype(np.float64)


def loss_and_grad_mixed(p, x, t, r16=_to_fp16, with_output=False):
    """Emulation of the training kernels: float16 weights and layer inputs
    in the forward (float32 sums, float32 bias, ReLU, float16 rounding of
    each activation), y and dL/dy in float32, then the backward through the
    float32 weights with float64 sums, ReLU masks from the float16
    activations, weight gradients from the float16 layer inputs.  Returns
    (mean loss, gradient) (and y with with_output).  r16 is the float16
    rounding (replaceable, so that without it the result can be held to the
    float64 gradient)."""
    p32 = np.asarray(p, np.float32)
    layers = unpack(p32)
    t = np.asarray(t, np.float32)
    n = x.shape[0]
    a = r16(np.asarray(x, np.float32))
    acts = [a]
    for li, (w, b) in enumerate(layers):
        w16 = r16(w)
        z = (a @ w16.T).astype(np.float32) + b.astype(np.float32)
        if li == 5:
            y = z.astype(np.float32)
            break
        a = r16(np.maximum(z, np.float32(0)))
        acts.append(a)
    w_ = (np.float32(1) / (y * y + np.float32(0.01))).astype(np.float32)
    d = (y - t).astype(np.float32)
    loss = float(np.sum((d.astype(np.float64) ** 2) * w_)) / (OUT * n)
    inv = np.float32(1) / np.float32(OUT * n)
    dz = (((np.float32(2) * d) * w_) * inv).astype(np.float64)
    grads = []
    for li in range(5, -1, -1):
        w, _ = layers[li]
        a_in = acts[li]
        grads.append((dz.T @ a_in, dz.sum(0)))
        if li > 0:
            dh = dz @ w.astype(np.float64)
            dz = dh * (acts[li] > 0)
    grads.reverse()
    g = np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])
    return (loss, g, y) if with_output else (loss, g)
