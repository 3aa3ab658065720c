"""CPU checks of the radiance-cache reference model (oracle/cache_ref.py):
the analytic gradient against central finite differences, the encoding's
first band against sin/cos and the recurrence against float64 sin(2^k pi v),
the float16 inference emulation against the float64 forward, and the
mixed-precision training emulation: without its float16 rounding it is the
float64 gradient, and its forward is the inference emulation."""
from __future__ import annotations

import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cache_ref as cr  # noqa: E402


def vsincos(orc):
    def f(x):
        s = np.empty(len(x), np.float32)
        c = np.empty(len(x), np.float32)
        for i, v in enumerate(x):
            s[i], c[i] = orc.sincos(float(v))
        return s, c
    return f


def small_params(rng, scale=1.0):
    p = np.zeros(cr.N_PARAMS, np.float64)
    off = 0
    for i, o in cr.LAYERS:
        p[off:off + o * i] = rng.uniform(-1, 1, o * i) * math.sqrt(6.0 / i) * scale
        off += o * i
        p[off:off + o] = rng.uniform(-0.1, 0.1, o)
        off += o
    return p


def test_param_count():
    assert cr.N_PARAMS == 64 * 288 + 64 + 4 * (64 * 64 + 64) + 3 * 64 + 3 == 35331


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = small_params(rng)
    x = rng.uniform(-1, 1, (16, cr.IN))
    t = rng.uniform(0, 2, (16, 3))
    loss, g = cr.loss_and_grad64(p, x, t)
    assert loss > 0
    # spot-check parameters in every layer (weights and biases)
    idx = []
    off = 0
    for i, o in cr.LAYERS:
        idx += list(rng.integers(off, off + o * i, 6)) + list(rng.integers(off + o * i, off + o * i + o, 2))
        off += o * i + o
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        # the stop-gradient weights are part of the loss definition: hold them fixed
        yp, _ = cr.forward64(pp, x)
        ym, _ = cr.forward64(pm, x)
        y0, _ = cr.forward64(p, x)
        w = 1.0 / (y0 * y0 + 0.01)
        lp = np.sum((yp - t) ** 2 * w) / (3 * len(x))
        lm = np.sum((ym - t) ** 2 * w) / (3 * len(x))
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_encoding_bands(orc):
    rng = np.random.default_rng(1)
    raw = rng.uniform(0, 1, (64, cr.RAW)).astype(np.float32)
    enc = cr.encode(raw, vsincos(orc))
    v = raw.astype(np.float64)
    assert np.allclose(enc[:, 0::24], np.sin(np.pi * v), atol=2e-7)
    assert np.allclose(enc[:, 1::24], np.cos(np.pi * v), atol=2e-7)
    for k in range(cr.BANDS):
        # the recurrence amplifies the phase and rounding error about 2.3x per band
        tol = 4e-7 * 2.5 ** k
        assert np.allclose(enc[:, 2 * k::24], np.sin(2.0 ** k * np.pi * v), atol=tol), k
        assert np.allclose(enc[:, 2 * k + 1::24], np.cos(2.0 ** k * np.pi * v), atol=tol), k


def test_fp16_emulation_close_to_fp64():
    rng = np.random.default_rng(2)
    p = small_params(rng).astype(np.float32)
    x = rng.uniform(-1, 1, (256, cr.IN)).astype(np.float32)
    y64, _ = cr.forward64(p, x)
    y16 = cr.forward_fp16(p, x)
    assert np.max(np.abs(y16 - y64)) <= 0.02 * np.mean(np.abs(y64)) + 1e-3


def test_mixed_training_emulation():
    rng = np.random.default_rng(8)
    p = small_params(rng)
    n = 64
    x = rng.uniform(-1, 1, (n, cr.IN)).astype(np.float32)
    t = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    # no rounding: the float64 gradient (up to float32 y and dL/dy)
    l0, g0 = cr.loss_and_grad_mixed(p, x, t, r16=lambda a: np.asarray(a, np.float64))
    l64, g64 = cr.loss_and_grad64(p.astype(np.float32).astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(l0 - l64) <= 1e-5 * l64
    assert np.linalg.norm(g0 - g64) <= 1e-5 * np.linalg.norm(g64)
    # the forward of the emulation is the inference emulation
    _, _, y = cr.loss_and_grad_mixed(p, x, t, with_output=True)
    assert np.array_equal(y, cr.forward_fp16(p, x))
    # the rounding matters, but only at float16 precision
    l1, g1 = cr.loss_and_grad_mixed(p, x, t)
    assert 0 < np.linalg.norm(g1 - g64) <= 3e-2 * np.linalg.norm(g64)
