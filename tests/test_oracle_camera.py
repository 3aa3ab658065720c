"""Pins the camera rays (generateRay.h, coordinates.h:1-33) by closed forms:
the pinhole ray through the image centre is the forward axis, all thin-lens
samples of a pixel meet on the focal plane, orthographic rays are parallel."""
import ctypes as C
import math

import numpy as np
import pytest

f32 = np.float32


def cam(orc, vtapi, model, W=64, H=48, lens_radius=5.0, focal=800.0):
    eye = np.array([300.0, 200.0, -900.0], np.float32)
    tgt = np.array([10.0, -20.0, 30.0], np.float32)
    fov = 0.7
    mv, inv_mv, p, inv_p = vtapi.camera_matrices(eye, tgt, fov, 0.1, 10000.0, model, W, H)
    c = orc.Camera()
    c.inv_mv[:] = list(inv_mv.reshape(-1))
    c.proj[:] = list(p.reshape(-1))
    c.inv_proj[:] = list(inv_p.reshape(-1))
    c.viewport[:] = [0, 0, W, H]
    c.win_z = 0.55
    c.lens_model = model
    c.lens_radius = lens_radius
    c.focal_distance = focal
    fwd = (tgt - eye) / np.linalg.norm(tgt - eye)
    return c, eye, fwd


def ray(orc, c, fx, fy, u):
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    orc.lib().orc_generate_ray(C.byref(c), fx, fy, np.asarray(u, np.float32), o, d)
    return o.astype(np.float64), d.astype(np.float64)


def test_pinhole_centre_ray_is_forward(orc, vtapi):
    c, eye, fwd = cam(orc, vtapi, 0)
    o, d = ray(orc, c, 32.0, 24.0, [0.5, 0.5, 0, 0])     # pixel (31.5+0.5, 23.5+0.5): jitter 0
    assert np.allclose(o, eye, atol=1e-3)
    assert np.allclose(d, fwd, atol=2e-6)
    assert np.linalg.norm(d) == pytest.approx(1.0, abs=1e-6)
    # a pixel to the right turns the ray towards +right
    o2, d2 = ray(orc, c, 60.0, 24.0, [0.5, 0.5, 0, 0])
    right = np.cross([0, 1, 0], fwd)
    right /= np.linalg.norm(right)
    assert d2 @ right > d @ right


def test_thin_lens_samples_converge_on_the_focal_plane(orc, vtapi):
    c, eye, fwd = cam(orc, vtapi, 1, lens_radius=25.0, focal=700.0)
    rng = np.random.default_rng(3)
    for px, py in [(10.5, 7.5), (40.5, 30.5)]:
        rays = [ray(orc, c, px, py, [a, b, 0, 0]) for a, b in rng.random((12, 2))]
        o0, d0 = ray(orc, c, px, py, [0.0, 0.0, 0, 0])   # lens centre
        # all rays pass through one point: take it from the lens-centre ray at the plane depth
        t_star = None
        pts = []
        for o, d in rays:
            # closest point between (o0,d0) and (o,d)
            w0 = o0 - o
            a, b, cc, dd, ee = d0 @ d0, d0 @ d, d @ d, d0 @ w0, d @ w0
            den = a * cc - b * b
            s = (b * ee - cc * dd) / den
            pts.append(o0 + s * d0)
        pts = np.array(pts)
        assert np.max(np.linalg.norm(pts - pts.mean(0), axis=1)) < 0.05
        # lens origins spread over the aperture, directions differ
        assert np.ptp([o for o, _ in rays], axis=0).max() > 5.0


def test_orthographic_rays_are_parallel(orc, vtapi):
    c, eye, fwd = cam(orc, vtapi, 2)
    o1, d1 = ray(orc, c, 3.5, 5.5, [0, 0, 0, 0])
    o2, d2 = ray(orc, c, 60.5, 40.5, [0, 0, 0, 0])
    assert np.allclose(d1, d2, atol=1e-7) and np.allclose(d1 / np.linalg.norm(d1), fwd, atol=1e-6)
    assert np.linalg.norm(o1 - o2) > 10.0
