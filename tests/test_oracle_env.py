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


# ---- generateImageFunction (image.cpp:285-346): resize, luminance, 3x3 blur ----

def gaussian3_weights():
    """make_kernel("gaussian", 3, 3): OIIO's Gaussian filter of width 3 is
    exp(-2 (x / 1.5)^2) inside |x| < 1.5, sampled at the pixel offsets -1, 0,
    1, the 2-D kernel the product of two, normalised to sum 1."""
    g = np.exp(-2.0 * (np.array([-1.0, 0.0, 1.0]) / 1.5) ** 2)
    k = np.outer(g, g)
    return k / k.sum()


def lanczos3(x):
    """Textbook Lanczos-3 window: sinc(x) sinc(x / 3) for |x| < 3."""
    x = np.asarray(x, np.float64)
    return np.where(np.abs(x) < 3.0, np.sinc(x) * np.sinc(x / 3.0), 0.0)


def test_image_function_interior_impulse_is_the_gaussian_kernel(orc):
    img = np.zeros((9, 12, 3), np.float32)
    img[4, 6] = 1.0
    f = orc.image_function(img)
    lum = (np.float32(0.2126) + np.float32(0.7152)) + np.float32(0.0722)
    expect = np.zeros((9, 12))
    expect[3:6, 5:8] = gaussian3_weights() * float(lum)
    assert f.shape == (9, 12)
    assert np.allclose(f, expect, rtol=0, atol=2e-7)
    # the 3x3 blur keeps the total (normalised kernel), and is symmetric
    assert abs(float(f.sum(dtype=np.float64)) - float(lum)) < 1e-6
    blk = f[3:6, 5:8]
    assert np.array_equal(blk, blk[::-1, :]) and np.array_equal(blk, blk[:, ::-1])


def test_image_function_edge_impulse_is_truncated_not_renormalised(orc):
    """Black outside the image: an impulse in the corner keeps only the
    kernel's four inside weights, so the total drops to their sum."""
    img = np.zeros((6, 7, 3), np.float32)
    img[0, 0] = (2.0, 2.0, 2.0)
    f = orc.image_function(img)
    lum = float((np.float32(2 * 0.2126) + np.float32(2 * 0.7152)) + np.float32(2 * 0.0722))
    k = gaussian3_weights()
    expect = np.zeros((6, 7))
    expect[0:2, 0:2] = k[1:3, 1:3] * lum
    assert np.allclose(f, expect, rtol=0, atol=4e-7)
    assert float(f.sum(dtype=np.float64)) == pytest.approx(k[1:, 1:].sum() * lum, rel=1e-6)
    assert float(f.sum(dtype=np.float64)) < 0.7 * lum


def test_image_function_resized_size_follows_the_reference(orc):
    # (unsigned)((float)w / max(w, h) * 512), image.cpp:313-314
    for (w, h) in [(1024, 512), (1000, 300), (300, 1000), (513, 7), (700, 699), (512, 512), (600, 600)]:
        img = np.ones((h, w, 3), np.float32)
        f = orc.image_function(img)
        m = max(w, h)
        if m > 512:
            ew = int(np.float32(np.float32(w) / np.float32(m)) * np.float32(512))
            eh = int(np.float32(np.float32(h) / np.float32(m)) * np.float32(512))
        else:
            ew, eh = w, h
        assert f.shape == (eh, ew), (w, h)


def test_image_function_constant_image_stays_constant_through_the_resize(orc):
    """Normalised resampling weights reproduce a constant (partition of
    unity): after the resize, the blurred luminance is the constant inside
    and the truncated-kernel sums at the border."""
    c = np.float32(3.5)
    img = np.full((512, 1024, 3), c, np.float32)
    f = orc.image_function(img)
    assert f.shape == (256, 512)
    lum = float((c * np.float32(0.2126) + c * np.float32(0.7152)) + c * np.float32(0.0722))
    assert np.allclose(f[1:-1, 1:-1], lum, rtol=2e-6)
    k = gaussian3_weights()
    assert np.allclose(f[0, 1:-1], k[1:, :].sum() * lum, rtol=2e-6)      # top row loses one kernel row
    assert f[0, 0] == pytest.approx(k[1:, 1:].sum() * lum, rel=2e-6)    # corner keeps a 2x2 block


def test_image_function_resize_is_lanczos3_in_destination_pixels(orc):
    """Halving an image: each destination pixel is the normalised Lanczos-3
    weighted sum of the source pixels whose centres lie within 3 destination
    pixels (6 source pixels) of its centre.  A linear ramp is reproduced
    exactly (symmetric taps), which fixes the half-pixel alignment; an
    impulse spreads over exactly the pixels of that support, with the
    kernel's negative lobes."""
    W, H = 1024, 40                      # resized to 512 x 20
    ramp = np.tile((np.arange(W, dtype=np.float32) + 0.5)[None, :, None], (H, 1, 3))
    lum_w = float((np.float32(0.2126) + np.float32(0.7152)) + np.float32(0.0722))
    f = orc.image_function(ramp)
    nw, nh = f.shape[1], f.shape[0]
    assert (nw, nh) == (512, 20)
    ratio = W / nw
    # undo the blur on interior rows: a ramp along x is unchanged by the
    # symmetric horizontal taps, and the rows are identical vertically
    centres = (np.arange(nw) + 0.5) * ratio
    inner = slice(8, nw - 8)
    assert np.allclose(f[5, inner], centres[inner] * lum_w, rtol=3e-6)

    img = np.zeros((H, W, 3), np.float32)
    p = 517
    img[:, p] = 1.0                      # a vertical line: a 1-D impulse along x
    g = orc.image_function(img, max_size=10 ** 6)
    assert g.shape == (H, W)             # no resize below the limit
    f = orc.image_function(img)
    # undo the horizontal blur of the interior row to read the resize
    row = f[9].astype(np.float64)
    kx = gaussian3_weights().sum(axis=0)          # horizontal weights of a vertically constant image
    # expected resized row by the textbook kernel, then the 1-D blur
    d = np.arange(nw)
    centre = (d + 0.5) * ratio
    taps = np.arange(-12, 13)
    r = np.zeros(nw)
    for x in d:
        src = np.floor(centre[x]).astype(int) + taps
        wts = lanczos3((src + 0.5 - centre[x]) / ratio)
        r[x] = wts[src == p].sum() / wts.sum()
    expect = np.convolve(r, kx[::-1], mode="same") * lum_w
    assert np.allclose(row, expect, atol=2e-6)
    support = np.abs((p + 0.5) - centre) < 3 * ratio
    assert np.all(np.abs(r[~support]) == 0) and np.count_nonzero(r) == support.sum()
    assert (r < 0).any()                          # Lanczos lobes


def test_image_function_luminance_commutes_with_the_resize(orc):
    """Resize and luminance are both linear, so taking the luminance first
    (a grey image of it) changes the result only by rounding."""
    rng = np.random.default_rng(5)
    img = (rng.random((300, 700, 3)) ** 3 * 50).astype(np.float32)
    f = orc.image_function(img)
    lum = ((img[..., 0] * np.float32(0.2126) + img[..., 1] * np.float32(0.7152)) + img[..., 2] * np.float32(0.0722))
    grey = np.repeat(lum[..., None], 3, axis=2) / np.float32(lum_sum := (np.float32(0.2126) + np.float32(0.7152)) + np.float32(0.0722))
    g = orc.image_function(grey.astype(np.float32))
    assert f.shape == g.shape == (219, 512)
    assert np.allclose(f, g * lum_sum, rtol=1e-5, atol=1e-5 * float(np.abs(f).max()))
