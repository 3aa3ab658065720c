"""Pins the environment CDF builder core (image.cpp:104-319), the CDF
sampling (envMapSample.h:23-123), the GL_CLAMP bilinear lookup and the
direction->uv mapping (coordinates.h:97-109) by closed forms."""
import ctypes as C
import math

import numpy as np
import pytest

f32 = np.float32


def env_with(orc, cu, cv, img=None):
    o = orc.OracleScene(np.full((2, 2, 2), -1, np.int32), np.zeros(7, np.float32))
    h = cu.shape[0]
    w = cu.shape[1] - 1
    if img is None:
        img = np.ones((h, w, 3), np.float32)
    o.set_environment(image=img, cdf_u=cu, cdf_v=cv, integral=1.0)
    return o


def test_uniform_function_image(orc):
    w, h = 64, 32
    cu, cv, integ = orc.calculate_cdf(np.ones((h, w), np.float32))
    x = np.arange(w + 1) / w
    assert np.allclose(cu, x[None, :], atol=2e-7) and np.all(cu[:, -1] == 1.0)
    s = np.sin(math.pi * (np.arange(h) + 0.5) / h)
    assert np.allclose(cv, np.r_[0, np.cumsum(s)] / s.sum(), atol=3e-7)
    assert integ == pytest.approx(s.mean() * 2 * math.pi ** 2, rel=2e-5)   # float running sum
    # uniform radiance -> pdf = luminance / integral ~ 1 / (4 pi)
    assert 1.0 / integ == pytest.approx(1 / (4 * math.pi), rel=2e-3)


def test_zero_function_image_falls_back_to_uniform(orc):
    w, h = 10, 7
    cu, cv, integ = orc.calculate_cdf(np.zeros((h, w), np.float32))
    assert integ == 0.0
    assert np.array_equal(cu, np.tile((np.arange(w + 1, dtype=np.float32) / f32(w)), (h, 1)))
    assert np.array_equal(cv, np.arange(h + 1, dtype=np.float32) / f32(h))


def test_one_hot_image_always_samples_its_texel(orc):
    w, h = 16, 8
    f = np.zeros((h, w), np.float32)
    f[5, 11] = 3.0
    cu, cv, _ = orc.calculate_cdf(f)
    o = env_with(orc, cu, cv)
    rng = np.random.default_rng(0)
    uv = np.zeros(2, np.float32)
    idx = np.zeros(2, np.int32)
    for su, sv in rng.uniform(1e-6, 1 - 1e-6, (500, 2)).astype(np.float32):
        orc.lib().orc_env_sample_uv(C.byref(o.env), su, sv, uv, idx)
        assert tuple(idx) == (11, 5)
        assert 11 / w <= uv[0] <= 12 / w and 5 / h <= uv[1] <= 6 / h


def test_binary_search_is_upper_bound(orc):
    rng = np.random.default_rng(1)
    w, h = 37, 19
    f = rng.random((h, w)).astype(np.float32) ** 4
    f[:, rng.integers(0, w, 8)] = 0.0       # zero-probability columns
    f[rng.integers(0, h, 3), :] = 0.0       # and rows
    cu, cv, _ = orc.calculate_cdf(f)
    o = env_with(orc, cu, cv)
    uv = np.zeros(2, np.float32)
    idx = np.zeros(2, np.int32)
    for su, sv in rng.random((3000, 2)).astype(np.float32):
        orc.lib().orc_env_sample_uv(C.byref(o.env), su, sv, uv, idx)
        row = np.searchsorted(cv, sv, side="right") - 1
        col = np.searchsorted(cu[row], su, side="right") - 1
        assert (idx[1], idx[0]) == (row, col)
        assert f[row, col] > 0 or cv[row + 1] == cv[row]


def test_bilinear_lookup_with_clamp_border(orc):
    w, h = 8, 4
    img = np.full((h, w, 3), 2.0, np.float32)
    cu, cv, _ = orc.calculate_cdf(np.ones((h, w), np.float32))
    o = env_with(orc, cu, cv, img)
    rgb = np.zeros(3, np.float32)
    tex = lambda u, v: (orc.lib().orc_env_texture(C.byref(o.env), u, v, rgb), rgb.copy())[1]
    assert np.allclose(tex(0.5, 0.5), 2.0)
    assert np.allclose(tex(0.0, 0.5), 1.0)            # half the black border
    assert np.allclose(tex(0.0, 0.0), 0.5)            # corner: a quarter of the image
    assert np.allclose(tex(-3.0, 0.5), 1.0)           # clamped coordinate
    img2 = np.zeros((h, w, 3), np.float32)
    img2[1, 2] = 4.0
    o2 = env_with(orc, cu, cv, img2)
    rgb2 = np.zeros(3, np.float32)
    orc.lib().orc_env_texture(C.byref(o2.env), (2 + 0.5) / w, (1 + 0.5) / h, rgb2)
    assert np.allclose(rgb2, 4.0)                     # texel centre reads the texel


def test_uv_from_direction(orc):
    rng = np.random.default_rng(2)
    uv = np.zeros(2, np.float32)
    for _ in range(2000):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        rot = float(rng.uniform(-3, 3))
        orc.lib().orc_uv_from_dir(d.astype(np.float32), rot, uv)
        u = ((math.atan2(d[2], d[0]) + math.pi + rot) / (2 * math.pi)) % 1.0
        v = math.acos(d[1]) / math.pi
        du = abs(uv[0] - u)
        assert min(du, 1 - du) < 2e-6 and abs(uv[1] - v) < 2e-6
