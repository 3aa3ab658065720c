"""Pins the oracle's integrator (pathTracer.fs main + directLighting, the
Lambert and microfacet BSDFs, the environment light) by closed forms and by
quadrature of the estimator's expectation."""
import math

import numpy as np
import pytest
from scipy import integrate

f32 = np.float32


def top_face_setup(orc, vtapi, mat, n=16, W=48, H=48):
    """One voxel in the top layer of an n^3 grid, a narrow camera looking
    down at its top face so every pixel's primary ray lands on that face."""
    off = np.full((n, n, n), -1, np.int32)
    x = z = n // 2
    off[z, n - 1, x] = 0
    o = orc.OracleScene(off, mat)
    c = o.bmin + (np.array([x, n - 1, z], np.float32) + f32(0.5)) * o.vsize
    c[1] = o.bmax[1]
    eye = (c + np.array([3.0, 300.0, 0.0], np.float32)).astype(np.float32)
    fov = 2 * math.atan(0.2 * o.vsize[0] / 300.0)
    mv, inv_mv, p, inv_p = vtapi.camera_matrices(eye, c, fov, 0.1, 10000.0, 0, W, H)
    o.set_camera(inv_mv, p, inv_p, np.array([0, 0, W, H], np.float32), 0.55, 0)
    o.set_environment((1, 1, 1), (1, 1, 1))
    o.set_settings(max_bounces=1, gamma=1.0)
    return o


def lambert_expectation(albedo):
    pL = 1 / (2 * math.pi)          # uniform hemisphere NEE pdf, lights.h:51
    pE = 1 / (4 * math.pi)          # environment evaluation pdf, lights.h:30

    def integrand(t):
        pB = math.cos(t) / math.pi
        wL = pL ** 2 / (pL ** 2 + pB ** 2)
        wB = pB ** 2 / (pB ** 2 + pE ** 2)
        return math.cos(t) * math.sin(t) * (wL + wB)

    val, _ = integrate.quad(integrand, 0, math.pi / 2)
    return 2 * albedo * val


def test_lambert_top_face_expectation(orc, vtapi):
    a = np.array([0.3, 0.55, 0.8], np.float32)
    mat = np.array([0, 0, 0, 0, *a], np.float32)
    o = top_face_setup(orc, vtapi, mat)
    W = H = 48
    frames = 96
    fh = o.first_hit(W, H, 7, 0)
    assert np.all(fh[..., 0] == 2) and np.all(fh[..., 6] == 1)     # every pixel hits the top face
    acc, cnt = o.render(W, H, 7, 0, frames)
    n = W * H * frames
    assert cnt["samples"] == n and cnt["interactions"] == n
    assert cnt["rays"] == 3 * n                                   # camera + shadow + continuation
    mean = acc[..., :3].reshape(-1, 3).mean(0)
    expect = np.array([lambert_expectation(float(v)) for v in a])
    # per-sample std is below albedo; 6 sigma of the mean
    tol = 6 * a / math.sqrt(n)
    assert np.all(np.abs(mean - expect) < tol), (mean, expect)
    # a dropped MIS term would move the mean by far more than the tolerance
    assert np.all(np.abs(expect - a) > 3 * tol)


def test_emission_only_at_first_vertex_and_exposure_gamma(orc, vtapi):
    """Emission is added once at bounce 0 (pathTracer.fs:241-245); exposure
    and gamma apply per sample (:294)."""
    Le = np.array([2.0, 1.0, 0.5], np.float32)
    mat = np.array([0, *Le, 0, 0, 0], np.float32)     # black emitter, no reflection
    o = top_face_setup(orc, vtapi, mat, W=8, H=8)
    o.set_settings(max_bounces=3, gamma=2.2, exposure=0.5)
    acc, _ = o.render(8, 8, 3, 0, 4)
    expect = np.power(Le.astype(np.float64) * 0.5, 1 / 2.2)
    assert np.allclose(acc[..., :3].reshape(-1, 3), expect, rtol=2e-6)


def test_background_gradient_on_miss(orc, vtapi):
    """A ray that misses the volume returns the gradient environment
    (lights.h:4-17): bottom*(1-max(0,y)) + top*max(0,y), tonemapped."""
    off = np.full((4, 4, 4), -1, np.int32)
    o = orc.OracleScene(off, np.array([0, 0, 0, 0, .5, .5, .5], np.float32))
    eye = np.array([0, 0, 5000], np.float32)
    tgt = np.array([0, 3000, 0], np.float32)
    W = H = 16
    mv, inv_mv, p, inv_p = vtapi.camera_matrices(eye, tgt, 0.3, 0.1, 10000.0, 0, W, H)
    o.set_camera(inv_mv, p, inv_p, np.array([0, 0, W, H], np.float32), 0.55, 0)
    top, bot = (0.2, 0.4, 0.9), (0.9, 0.5, 0.1)
    o.set_environment(top, bot)
    o.set_settings(max_bounces=2, gamma=1.0)
    acc, cnt = o.render(W, H, 1, 0, 1)
    assert cnt["interactions"] == 0
    d = tgt - eye
    d = d / np.linalg.norm(d)
    bias = max(0.0, d[1])
    expect = np.array(bot) * (1 - bias) + np.array(top) * bias
    centre = acc[H // 2, W // 2, :3]
    assert np.allclose(centre, expect, atol=2e-2)


def test_metal_sampling_matches_its_pdf(orc, vtapi):
    """Microfacet sampling (microfacet.h:41-61): with NEE switched off (env
    importance sampling is only used for image environments, so here the
    shadow-ray contribution is computed by quadrature too), the estimator's
    mean equals the quadrature of its expectation over the Blinn half-vector
    density, which checks that sampled directions follow the pdf the code
    reports."""
    refl = np.array([0.9, 0.6, 0.3], np.float32)
    e = 0.35                                                    # exponent = roughness (metal.h:14)
    mat = np.array([1, 0, 0, 0, *refl, e], np.float32)
    o = top_face_setup(orc, vtapi, mat, W=32, H=32)
    frames = 160
    acc, cnt = o.render(32, 32, 11, 0, frames)
    n = 32 * 32 * frames
    got = acc[..., :3].reshape(-1, 3).astype(np.float64)
    # wo: the view direction of the (nearly vertical) camera, in the face's local frame
    fh_dir = np.array([3.0, 300.0, 0.0])
    wo = fh_dir / np.linalg.norm(fh_dir)
    wo_l = np.array([0.0, wo[1], math.hypot(wo[0], wo[2])])    # azimuth is irrelevant below

    inv2pi, twopi = 1 / (2 * math.pi), 2 * math.pi
    ior = 7.3
    r0 = ((1 - ior) / (1 + ior)) ** 2

    def G(wo, wi, wh):
        d = abs(wo @ wh)
        return min(1.0, 2 * wh[1] * wo[1] / d, 2 * wh[1] * wi[1] / d)

    def F(c):
        return r0 + (1 - r0) * (1 - c) ** 5

    def D_f(wh):
        return (e + 2) * inv2pi * wh[1] ** e

    def D_pdf(wo, wh):
        d = wo @ wh
        return 0.0 if d <= 0 else (e + 1) * wh[1] ** e / (twopi * 4 * d)

    pL = 1 / (2 * math.pi)
    pE = 1 / (4 * math.pi)

    def bsdf_term(ct, ph):
        st = math.sqrt(max(0.0, 1 - ct * ct))
        wh = np.array([st * math.cos(ph), ct, st * math.sin(ph)])
        wi = -wo_l + 2 * (wo_l @ wh) * wh
        pdf = D_pdf(wo_l, wh)
        if wi[1] <= 0 or pdf <= 0:
            return None
        wh2 = (wi + wo_l) / np.linalg.norm(wi + wo_l)
        f = D_f(wh) * G(wo_l, wi, wh2) * F(wi @ wh2) / (4 * wo_l[1] * wi[1])   # times refl^2 per channel
        w = pdf ** 2 / (pdf ** 2 + pE ** 2)
        return f * abs(wi[1]) / pdf * w

    def nee_term(wi):
        wh = (wi + wo_l) / np.linalg.norm(wi + wo_l)
        f = D_f(wh) * G(wo_l, wi, wh) * F(wi @ wh) / (4 * wo_l[1] * wi[1])     # times refl
        pdf = D_pdf(wo_l, wh)
        w = pL ** 2 / (pL ** 2 + pdf ** 2)
        return f * wi[1] * w / pL

    # BSDF sampling: cos(theta_h) = u^(1/(e+1)), phi uniform -> expectation over u, phi
    bs, _ = integrate.dblquad(lambda ph, u: (bsdf_term(u ** (1 / (e + 1)), ph) or 0.0), 0, 1, 0, 2 * math.pi,
                              epsabs=1e-6)
    bs /= 2 * math.pi
    # NEE: uniform hemisphere y = u, phi uniform
    ne, _ = integrate.dblquad(lambda ph, y: nee_term(np.array([math.sqrt(max(0, 1 - y * y)) * math.cos(ph), y,
                                                                math.sqrt(max(0, 1 - y * y)) * math.sin(ph)])),
                              0, 1, 0, 2 * math.pi, epsabs=1e-6)
    ne /= 2 * math.pi
    expect = refl.astype(np.float64) ** 2 * bs + refl.astype(np.float64) * ne
    mean = got.mean(0)
    sd = got.std(0) / math.sqrt(n)
    assert np.all(np.abs(mean - expect) < 6 * sd + 1e-3), (mean, expect, sd)


def test_plastic_is_metal_of_unit_reflectance(orc, vtapi):
    """plastic.h: the microfacet BSDF with reflectance 1, whatever diffuse
    albedo the record stores (material.h:28-33). A plastic face renders
    bit-identically to a metal face of reflectance 1 and the same exponent,
    and differently from a metal face of another reflectance or exponent,
    so neither the stored albedo nor a misplaced exponent goes unnoticed."""
    e = 6.0
    plastic = np.array([2, 0, 0, 0, 0.8, 0.2, 0.1, e], np.float32)
    metal1 = np.array([1, 0, 0, 0, 1.0, 1.0, 1.0, e], np.float32)
    renders = []
    for mat in (plastic, metal1, np.array([1, 0, 0, 0, 0.8, 0.2, 0.1, e], np.float32),
                np.array([1, 0, 0, 0, 1.0, 1.0, 1.0, 2 * e], np.float32)):
        o = top_face_setup(orc, vtapi, mat, W=24, H=24)
        acc, _ = o.render(24, 24, 5, 0, 8)
        renders.append(acc)
    assert np.isfinite(renders[0]).any()
    assert np.array_equal(renders[0].view(np.uint32), renders[1].view(np.uint32))
    assert not np.array_equal(renders[0].view(np.uint32), renders[2].view(np.uint32))
    assert not np.array_equal(renders[0].view(np.uint32), renders[3].view(np.uint32))
