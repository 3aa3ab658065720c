"""GPU parity of the renderer object (voxeltoy::Renderer through vtr_*): its
progressive passes, .vox and mesh loading, picking services and image
output against the CPU oracle on the same inputs.  Radiance is compared
bit for bit (relative L2 <= 1e-4 asserted alongside); the focal distance of
the picking service is checked against float64 geometry at rel 1e-5."""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.array_equal(np.isfinite(a), np.isfinite(b)), "finiteness differs"
    f = np.isfinite(a)
    return np.linalg.norm((a - b)[f]) / max(np.linalg.norm(b[f]), 1e-30)


def oracle_of(orc, r, bounces, selected=(0, 0, 0), top=(1, 1, 1), bottom=(1, 1, 1)):
    """Oracle scene fed with the renderer's own host copies and camera block."""
    off, bmin, bmax = r.volume()
    o = orc.OracleScene(off, r.material_data(), r.emissive(), bmin, bmax)
    cb = r.camera_block()
    o.set_camera(np.array(cb.inv_modelview, np.float32), np.array(cb.proj, np.float32),
                 np.array(cb.inv_proj, np.float32), np.array(cb.viewport, np.float32), cb.win_z, cb.lens_model,
                 cb.lens_radius, cb.focal_distance)
    o.set_environment(top, bottom)
    o.set_settings(max_bounces=bounces, selected=selected)
    return o


def make_renderer(vr, w, h, bounces=2, top=(1, 1, 1), bottom=(1, 1, 1)):
    r = vr.Renderer(0)
    r.initialize("")
    r.resize_frame(w, h)
    s = r.render_settings()
    s.max_bounces = bounces
    s.background_top[:] = list(top)
    s.background_bottom[:] = list(bottom)
    r.update_render_settings(s)
    return r


@pytest.fixture(scope="module")
def vr():
    from voxeltoy_amd import renderer
    return renderer


def small_volume(n=32, seed=5):
    from voxeltoy_amd import scenes
    sc = scenes.small_scene(seed, n=n)
    return sc.offsets(), sc.mat


def test_renderer_passes_bit_exact(vr, orc):
    offs, mat = small_volume()
    r = make_renderer(vr, 64, 48)
    r.load_volume(offs, mat)
    r.focus_on_bounds()
    r.orbit(math.radians(115), math.radians(40))
    r.set_seed(11)
    for k in range(3):
        assert r.render() == vr.RR_SAMPLES_PENDING
    assert r.number_samples() == 3
    got = r.read_accum()
    s = r.render_settings()
    o = oracle_of(orc, r, 2)
    ref, _ = o.render(64, 48, 11, 0, 3, rng_frame_max=s.max_samples - 1)
    assert rel_l2(got, ref) <= 1e-4
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_checkpoint_resume_equals_uninterrupted_render(vr, tmp_path):
    """3 passes, checkpoint, a new renderer with the same scene and camera
    loads it and renders 2 more: the image equals 5 uninterrupted passes;
    the same through the context ABI with FP64 sums.  A checkpoint of
    another frame size is refused."""
    from voxeltoy_amd import api, workloads
    offs, mat = small_volume()

    def fresh():
        r = make_renderer(vr, 64, 48)
        r.load_volume(offs, mat)
        r.focus_on_bounds()
        r.orbit(math.radians(115), math.radians(40))
        return r

    ref = fresh()
    ref.set_seed(5)
    for _ in range(5):
        ref.render()
    a = fresh()
    a.set_seed(5)
    for _ in range(3):
        a.render()
    path = str(tmp_path / "progress.ckpt")
    a.save_checkpoint(path)
    b = fresh()
    b.set_seed(999)                       # the checkpoint carries the seed
    b.load_checkpoint(path)
    assert b.number_samples() == 3
    for _ in range(2):
        b.render()
    assert b.number_samples() == 5
    assert np.array_equal(b.read_accum().view(np.uint32), ref.read_accum().view(np.uint32))
    c = make_renderer(vr, 32, 48)
    c.load_volume(offs, mat)
    with pytest.raises(vr.RendererError):
        c.load_checkpoint(path)
    # the context ABI: running mean and FP64 sums written back
    wl = workloads.small(1, width=48, height=40)
    x = api.Context(0)
    workloads.load(x, wl, fp64=True)
    x.render(wl.seed, 0, 4)
    whole, whole64 = x.read_accum(), x.read_sum64()
    y = api.Context(0)
    workloads.load(y, wl, fp64=True)
    y.render(wl.seed, 0, 2)
    acc, s64 = y.read_accum(), y.read_sum64()
    z = api.Context(0)
    workloads.load(z, wl, fp64=True)
    z.write_accum(acc)
    z.write_sum64(s64)
    z.render(wl.seed, 2, 2)
    assert np.array_equal(z.read_accum().view(np.uint32), whole.view(np.uint32))
    assert np.array_equal(z.read_sum64().view(np.uint64), whole64.view(np.uint64))
    for ctx in (x, y, z):
        ctx.close()


def test_renderer_finishes_at_max_samples(vr):
    offs, mat = small_volume(16)
    r = make_renderer(vr, 16, 16, bounces=1)
    r.load_volume(offs, mat)
    s = r.render_settings()
    s.max_samples = 2
    r.update_render_settings(s)
    # reference: pending while the pre-increment count is below max_samples
    assert [r.render() for _ in range(4)] == [0, 0, 1, 1]


def test_vox_fixture_renders_bit_exact(vr, orc):
    r = make_renderer(vr, 80, 60, bounces=2, top=(2.0, 1.8, 1.6), bottom=(0.3, 0.25, 0.2))
    r.load_vox_file(os.path.join(GOLDEN, "scene_fall.vox"))
    r.focus_on_bounds()
    r.orbit(math.radians(120), math.radians(200))
    r.set_seed(3)
    r.render()
    r.render()
    got = r.read_accum()
    o = oracle_of(orc, r, 2, top=(2.0, 1.8, 1.6), bottom=(0.3, 0.25, 0.2))
    ref, _ = o.render(80, 60, 3, 0, 2, rng_frame_max=r.render_settings().max_samples - 1)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_mesh_loading_voxelizes_like_the_oracle(vr, orc):
    from voxeltoy_amd import api
    path = os.path.join(GOLDEN, "bunny.obj")
    v, t, bmin, bmax = vr.load_obj(path)
    r = make_renderer(vr, 32, 32)
    r.load_mesh(path)
    off, _, _ = r.volume()
    assert off.shape == (64, 64, 64)
    M = api.mesh_transform(bmin, bmax, [64] * 3)
    ref = orc.voxelize(v, t, M, [64] * 3, mode=0)
    assert ref.sum() > 1000
    assert np.array_equal(off >= 0, ref > 0)
    assert set(np.unique(off)) == {-1, 0}
    # and the device grid is the host copy
    from voxeltoy_amd.api import lib
    dev = np.zeros_like(off)
    assert lib().vt_read_volume(r.context_handle(), dev.reshape(-1)) == 0
    assert np.array_equal(dev, off)


def block_scene():
    """An 8^3 block in the middle of a 32^3 grid, two materials."""
    n = 32
    offs = -np.ones((n, n, n), np.int32)
    offs[12:20, 12:20, 12:20] = 0
    offs[12:20, 12:20, 12:16] = 7
    mat = np.array([0, 0, 0, 0, .6, .5, .4, 0, 0, 0, 0, .2, .7, .3], np.float32)
    return offs, mat


def centre_ray(r):
    c = r.camera()
    d = (c["target"] - c["eye"]).astype(np.float64)
    return c["eye"].astype(np.float64), d / np.linalg.norm(d)


def slab(o, d, lo, hi):
    with np.errstate(divide="ignore"):
        t0, t1 = (lo - o) / d, (hi - o) / d
    tmin, tmax = np.max(np.minimum(t0, t1)), np.min(np.maximum(t0, t1))
    return None if tmax < 0 or tmin > tmax else max(0.0, tmin)


def test_picking_focal_distance_and_active_voxel(vr, orc):
    offs, mat = block_scene()
    r = make_renderer(vr, 64, 64, bounces=1)
    r.load_volume(offs, mat)
    r.set_eye_target([150, 260, -900], [0, 0, 0])
    # image centre: NDC (0, 0) is the forward axis
    r.request_action(0.5, 0.5, action=vr.PA_SELECT_FOCAL_POINT)
    r.request_action(0.5, 0.5, action=vr.PA_SELECT_ACTIVE_VOXEL)
    r.render()
    o, d = centre_ray(r)
    _, bmin, bmax = r.volume()
    t = slab(o, d, bmin.astype(np.float64), bmax.astype(np.float64))
    assert t is not None
    orc_s = orc.OracleScene(offs, mat, None, bmin, bmax)
    rec = orc_s.trace_rays((o + t * d).astype(np.float32)[None], d.astype(np.float32)[None])[0]
    assert rec[0] == 2
    vs = (bmax - bmin) / 32
    lo = rec[1:4] * vs + bmin
    want = slab(o, d, lo.astype(np.float64), (lo + vs).astype(np.float64))
    assert math.isclose(r.focal_distance(), want, rel_tol=1e-5)
    assert np.array_equal(r.selected_voxel(), rec[1:4])
    # the accumulated pass used the new focal distance and paints the selection red
    cb = r.camera_block()
    assert cb.focal_distance == np.float32(r.focal_distance())
    got = r.read_accum()
    ref, _ = oracle_of(orc, r, 1, selected=tuple(rec[1:4])).render(64, 64, 0, 0, 1,
                                                                     rng_frame_max=r.render_settings().max_samples - 1)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_picking_miss_resets_selection(vr):
    offs, mat = block_scene()
    r = make_renderer(vr, 32, 32, bounces=1)
    r.load_volume(offs, mat)
    r.set_eye_target([0, 0, -900], [0, 5000, -800])   # looking up and away from the volume
    r.request_action(0.5, 0.5, action=vr.PA_SELECT_ACTIVE_VOXEL)
    r.request_action(0.5, 0.5, action=vr.PA_SELECT_FOCAL_POINT)
    r.render()
    assert np.array_equal(r.selected_voxel(), [0, 0, 0])
    assert r.focal_distance() == pytest.approx(99999999.0)


def test_add_and_remove_voxel(vr, orc):
    offs, mat = block_scene()
    r = make_renderer(vr, 32, 32, bounces=1)
    r.load_volume(offs, mat)
    # looking down +Z at the block's -Z face, off the voxel boundaries:
    # world x = -40 is voxel 14, y = 13 is voxel 16
    r.set_eye_target([-40, 13, -900], [-40, 13, 0])
    r.request_action(0.5, 0.5, action=vr.PA_SELECT_ACTIVE_VOXEL)
    r.render()
    sel = r.selected_voxel()
    assert tuple(sel) == (14, 16, 12) and offs[12, 16, 14] == 7
    r.request_action(0.5, 0.5, action=vr.PA_ADD_VOXEL)   # no motion: along the selected face normal (-Z)
    r.render()
    off, _, _ = r.volume()
    assert off[11, sel[1], sel[0]] == 7
    assert (off != offs).sum() == 1
    r.request_action(0.5, 0.5, action=vr.PA_REMOVE_VOXEL)
    r.render()
    off2, _, _ = r.volume()
    assert off2[sel[2], sel[1], sel[0]] == -1
    # the device grid follows the host copy
    from voxeltoy_amd.api import lib
    dev = np.zeros_like(off2)
    assert lib().vt_read_volume(r.context_handle(), dev.reshape(-1)) == 0
    assert np.array_equal(dev, off2)


def test_save_image_pfm_and_ppm(vr, tmp_path):
    offs, mat = small_volume(16)
    r = make_renderer(vr, 24, 16, bounces=1)
    r.load_volume(offs, mat)
    r.focus_on_bounds()
    r.render()
    acc = r.read_accum()
    r.save_image(str(tmp_path / "a.pfm"))
    raw = (tmp_path / "a.pfm").read_bytes()
    head = raw.split(b"\n", 3)
    assert head[0] == b"PF" and head[1] == b"24 16" and float(head[2]) < 0
    px = np.frombuffer(head[3], "<f4").reshape(16, 24, 3)
    assert np.array_equal(px, acc[..., :3])            # PFM rows are bottom-up, like the buffer
    r.save_image(str(tmp_path / "a.ppm"))
    ppm = (tmp_path / "a.ppm").read_bytes().split(b"\n", 3)
    img = np.frombuffer(ppm[3], np.uint8).reshape(16, 24, 3)
    want = np.round(np.clip(np.nan_to_num(acc[::-1, :, :3]), 0, 1) * 255).astype(np.uint8)
    assert np.abs(img.astype(int) - want.astype(int)).max() <= 1
    with pytest.raises(vr.RendererError):
        r.save_image(str(tmp_path / "a.exr"))


def test_hdr_background_renders_like_the_same_pfm(vr, tmp_path):
    """An environment image given as Radiance RGBE (.hdr) renders exactly
    like its decoded floats given as PFM: the renderer reads both formats
    into the same environment (image, CDF and all)."""
    rng = np.random.default_rng(3)
    h, w = 32, 64
    sky = (rng.uniform(0.05, 2.0, (h, w, 3)) * np.linspace(2.0, 0.2, h)[:, None, None]).astype(np.float64)
    sky[4:6, 40:43] = 900.0                                          # a sun
    mx = sky.max(axis=-1)
    m, e = np.frexp(mx)
    rgbe = np.zeros((h, w, 4), np.uint8)
    rgbe[..., :3] = np.floor(sky * (m * 256.0 / mx)[..., None]).astype(np.uint8)
    rgbe[..., 3] = (e + 128).astype(np.uint8)
    # what the format's reader decodes: (r, g, b) * 2^(e - 136), in float
    decoded = (rgbe[..., :3].astype(np.float32) * np.ldexp(np.float32(1.0), rgbe[..., 3].astype(np.int32) - 136)[..., None]).astype(np.float32)
    hdr = tmp_path / "sky.hdr"
    hdr.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w) + rgbe.tobytes())
    pfm = tmp_path / "sky.pfm"
    pfm.write_bytes(b"PF\n%d %d\n-1.0\n" % (w, h) + decoded[::-1].astype("<f4").tobytes())   # PFM: bottom row first
    offs, mat = small_volume(24)
    imgs = []
    for path in (hdr, pfm):
        r = make_renderer(vr, 40, 30, bounces=2)
        r.load_volume(offs, mat)
        r.focus_on_bounds()
        s = r.render_settings()
        s.background_image = str(path).encode()
        r.update_render_settings(s)
        r.set_seed(7)
        for _ in range(2):
            r.render()
        imgs.append(r.read_accum())
    assert np.isfinite(imgs[0]).any() and imgs[0][..., :3].max() > 0
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
