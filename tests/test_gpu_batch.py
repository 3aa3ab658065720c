"""GPU: batched multi-scene rendering and the autotuner.  A batch must give
every scene exactly the image of its own single-scene render under the same
PCG streams (bit for bit), with and without the recorded HIP graph and on
graph replays with new seeds; the autotuner may only change speed."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mixed_scenes():
    """Six small scenes of different sizes, resolutions, spp, bounces,
    environments and lens models."""
    from voxeltoy_amd import scenes, workloads
    out = []
    specs = [(16, 24, 16, 3, 1, 0), (32, 40, 32, 2, 2, 1), (24, 16, 48, 1, 0, 0), (48, 64, 40, 2, 3, 2),
             (20, 32, 32, 4, 1, 1), (40, 48, 24, 1, 2, 0)]
    for i, (n, W, H, spp, bounces, lens) in enumerate(specs):
        sc = scenes.small_scene(10 + i, n=n)
        bmin, bmax = workloads.reference_bounds(sc.shape)
        cam = scenes.orbit_camera(bmin, bmax, math.radians(110 + 5 * i), math.radians(30 * i), W, H,
                                  lens_model=lens, lens_radius=25.0 if lens == 1 else 0.0,
                                  focal_distance=1200.0)
        wl = workloads.Workload(f"s{i}", sc, cam, W, H, spp=spp, bounces=bounces, seed=100 + i, bmin=bmin,
                                bmax=bmax, env_top=(1.5, 1.4, 1.2), env_bottom=(0.3, 0.3, 0.35))
        if i == 3:
            img = scenes.sky_environment(7, 64, 32, sun_peak=50.0, sun_exponent=200.0)
            from voxeltoy_amd import api
            wl.env_image, wl.cdf = img, api.env_build_cdf(img)
        out.append(wl)
    return out


def singles(api, workloads, wls, seeds, n0):
    imgs = []
    for wl, sd, z in zip(wls, seeds, n0):
        c = api.Context(0)
        workloads.load(c, wl)
        c.render(sd, z, wl.spp)
        imgs.append(c.read_accum())
        c.close()
    return imgs


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_scene_renders(graph):
    from voxeltoy_amd import api, workloads
    wls = mixed_scenes()
    seeds = [wl.seed for wl in wls]
    n0 = [0, 0, 2, 0, 1, 0]
    ref = singles(api, workloads, wls, seeds, n0)
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    b = api.Batch(ctxs)
    b.render(seeds, [wl.spp for wl in wls], n0=n0, graph=graph)
    for i, c in enumerate(ctxs):
        got = c.read_accum()
        assert np.array_equal(got.view(np.uint32), ref[i].view(np.uint32)), f"scene {i} differs"
    if graph:
        # replay of the recorded graph with new seeds (same shape) after clearing
        seeds2 = [s + 7 for s in seeds]
        for c in ctxs:
            c.reset_accum()
        b.render(seeds2, [wl.spp for wl in wls], n0=n0, graph=True)
        ref2 = singles(api, workloads, wls, seeds2, n0)
        for i, c in enumerate(ctxs):
            assert np.array_equal(c.read_accum().view(np.uint32), ref2[i].view(np.uint32)), f"replay scene {i}"
    b.close()
    for c in ctxs:
        c.close()


def long_scene(nx, seed):
    """A 6 x 5 x nx grid seen down its length (floor, sparse boxes, emitters
    at the far end), 48 x 12 pixels, 2 spp, 2 bounces."""
    from voxeltoy_amd import api, scenes, workloads
    nz, ny = 6, 5
    rng = np.random.default_rng(seed)
    occ = rng.random((nz, ny, nx)) < 0.001
    occ[:, 0, :] = True
    recs = [scenes.lambert((0, 0, 0), (0.6, 0.6, 0.6)), scenes.lambert((6.0, 5.0, 4.0), (0.5, 0.5, 0.5))]
    mat, offs = scenes.build_materials(recs)
    id_offset = np.full(256, -1, np.int32)
    id_offset[1:3] = offs
    ids = np.where(occ, np.uint8(1), np.uint8(0))
    ids[:, 1:3, nx - 8:] = 2
    sc = scenes.Scene(ids, id_offset, mat)
    sc.emissive = api.prune_emissive(sc.offsets(), scenes.emissive_list(ids, id_offset, mat))
    bmin, bmax = workloads.reference_bounds(sc.shape)
    eye = np.array([bmin[0] - 20.0, 0.8 * bmax[1], 0.1 * bmax[2]], np.float32)
    cam = scenes.CameraSetup(eye, np.array([bmax[0], 0.3 * bmax[1], 0.0], np.float32), 0.06)
    return workloads.Workload(f"long{nx}", sc, cam, 48, 12, spp=2, bounces=2, seed=seed, bmin=bmin, bmax=bmax,
                              env_top=(1.2, 1.2, 1.3), env_bottom=(0.2, 0.2, 0.2))


@pytest.mark.parametrize("nx", [1016, 1017])
def test_batch_with_a_long_grid_equals_single_renders(nx):
    """The merged launch marches with packed coordinates only when every
    merged scene's extents are at most 1016: a 1016-long grid keeps that
    loop, a 1017-long one switches the whole merged launch to the unpacked
    loop.  Either way each scene equals its own render."""
    from voxeltoy_amd import api, workloads
    wls = mixed_scenes()[:3] + [long_scene(nx, 41)]
    seeds = [wl.seed for wl in wls]
    n0 = [0] * len(wls)
    ref = singles(api, workloads, wls, seeds, n0)
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    b = api.Batch(ctxs)
    b.render(seeds, [wl.spp for wl in wls], n0=n0, graph=True)
    for i, c in enumerate(ctxs):
        assert np.array_equal(c.read_accum().view(np.uint32), ref[i].view(np.uint32)), f"scene {i} differs"
    b.close()
    for c in ctxs:
        c.close()


def test_batch_chunked_and_stats():
    """Path capacity smaller than the batch: chunks cut across scenes."""
    from voxeltoy_amd import api, workloads
    wls = mixed_scenes()[:4]
    seeds = [wl.seed for wl in wls]
    ref = singles(api, workloads, wls, seeds, [0] * 4)
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    ctxs[0].set_render_mode(1, 3001)
    b = api.Batch(ctxs)
    b.render(seeds, [wl.spp for wl in wls], stats=True)
    for i, c in enumerate(ctxs):
        assert np.array_equal(c.read_accum().view(np.uint32), ref[i].view(np.uint32)), f"scene {i} differs"
    st = b.stats()
    assert st["samples"] == sum(wl.width * wl.height * wl.spp for wl in wls)
    assert st["rays"] >= st["samples"] and st["kernel_ms"] > 0
    b.close()
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batch_with_scenes_of_their_own_launches(graph):
    """Scenes of 2^22 paths or more take their own launches (on their own
    streams), the rest one shared launch; both kinds equal their single
    renders, also on a graph replay with new seeds, and the counters add up
    over both kinds."""
    from voxeltoy_amd import api, scenes, workloads
    wls = mixed_scenes()[:3]
    big = []
    for i, (n, W, H, spp) in enumerate([(32, 1024, 1024, 4), (40, 1024, 768, 6)]):
        sc = scenes.small_scene(30 + i, n=n)
        bmin, bmax = workloads.reference_bounds(sc.shape)
        cam = scenes.orbit_camera(bmin, bmax, math.radians(120), math.radians(40 + 50 * i), W, H)
        big.append(workloads.Workload(f"b{i}", sc, cam, W, H, spp=spp, bounces=2, seed=300 + i, bmin=bmin, bmax=bmax,
                                      env_top=(1.5, 1.4, 1.2), env_bottom=(0.3, 0.3, 0.35)))
    wls = [wls[0], big[0], wls[1], big[1], wls[2]]
    assert sum(wl.width * wl.height * wl.spp >= (1 << 22) for wl in wls) == 2
    seeds = [wl.seed for wl in wls]
    ref = singles(api, workloads, wls, seeds, [0] * 5)
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    b = api.Batch(ctxs)
    b.render(seeds, [wl.spp for wl in wls], graph=graph, stats=True)
    for i, c in enumerate(ctxs):
        assert np.array_equal(c.read_accum().view(np.uint32), ref[i].view(np.uint32)), f"scene {i} differs"
    assert b.stats()["samples"] == sum(wl.width * wl.height * wl.spp for wl in wls)
    seeds2 = [s + 11 for s in seeds]
    for c in ctxs:
        c.reset_accum()
    b.render(seeds2, [wl.spp for wl in wls], graph=graph)
    ref2 = singles(api, workloads, wls, seeds2, [0] * 5)
    for i, c in enumerate(ctxs):
        assert np.array_equal(c.read_accum().view(np.uint32), ref2[i].view(np.uint32)), f"second render scene {i}"
    b.close()
    for c in ctxs:
        c.close()


def test_batch_rejects_bad_input():
    from voxeltoy_amd import api
    c = api.Context(0)
    b = api.Batch([c])
    with pytest.raises(api.VTError):
        b.render([1], [1])          # no volume / frame yet
    b.close()
    with pytest.raises(api.VTError):
        api.Batch([])
    d = api.Context(0)
    with pytest.raises(api.VTError, match="more than once"):
        api.Batch([c, d, c])        # one frame would take two scenes' samples
    api.Batch([c, d]).close()
    d.close()
    c.close()


def test_autotune_keeps_image_and_accumulation():
    from voxeltoy_amd import api, workloads
    wl = mixed_scenes()[1]
    c = api.Context(0)
    workloads.load(c, wl)
    c.render(wl.seed, 0, 1)
    before = c.read_accum()
    t = c.autotune(passes=2)
    assert t["candidates"] >= 10 and t["render_mode"] in (0, 1) and t["ms"] > 0
    assert np.array_equal(c.read_accum().view(np.uint32), before.view(np.uint32))
    c.render(wl.seed, 1, 2)
    tuned = c.read_accum()
    d = api.Context(0)
    workloads.load(d, wl)
    d.render(wl.seed, 0, 1)
    d.render(wl.seed, 1, 2)
    assert np.array_equal(tuned.view(np.uint32), d.read_accum().view(np.uint32))
    c.close()
    d.close()


def _big_and_small():
    from voxeltoy_amd import scenes, workloads
    wls = mixed_scenes()[:3]
    big = []
    for i, (n, W, H, spp) in enumerate([(32, 1024, 1024, 4), (40, 1024, 768, 6)]):
        sc = scenes.small_scene(30 + i, n=n)
        bmin, bmax = workloads.reference_bounds(sc.shape)
        cam = scenes.orbit_camera(bmin, bmax, math.radians(120), math.radians(40 + 50 * i), W, H)
        big.append(workloads.Workload(f"b{i}", sc, cam, W, H, spp=spp, bounces=2, seed=300 + i, bmin=bmin, bmax=bmax,
                                      env_top=(1.5, 1.4, 1.2), env_bottom=(0.3, 0.3, 0.35)))
    return [wls[0], big[0], wls[1], big[1], wls[2]]


def test_batch_autotune_keeps_images_and_accumulations():
    """vt_autotune_batch times graph replays over the solo threshold, the
    path capacity and the merged launch's tuning, restores every scene's
    accumulation, and the batch rendered with the chosen tuning (solo scenes
    inside the graph) still equals each scene's single render."""
    from voxeltoy_amd import api, workloads
    wls = _big_and_small()
    seeds = [wl.seed for wl in wls]
    counts = [wl.spp for wl in wls]
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    b = api.Batch(ctxs)
    b.render(seeds, [1] * len(wls), graph=True)
    before = [c.read_accum() for c in ctxs]
    t = b.autotune(seeds, counts)
    print("batch tuning", t, flush=True)
    assert t["candidates"] == 17 and t["ms"] > 0
    assert t["solo_paths"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24, (1 << 64) - 1)
    assert t["max_paths"] in (4 << 20, 8 << 20, 16 << 20)
    assert b.tuning()["solo_paths"] == t["solo_paths"]
    for i, c in enumerate(ctxs):
        assert np.array_equal(c.read_accum().view(np.uint32), before[i].view(np.uint32)), f"scene {i} not restored"
    ref = singles(api, workloads, wls, seeds, [0] * len(wls))
    for c in ctxs:
        c.reset_accum()
    b.render(seeds, counts, graph=True)
    for i, c in enumerate(ctxs):
        assert np.array_equal(c.read_accum().view(np.uint32), ref[i].view(np.uint32)), f"scene {i} differs"
    # every scene solo, then none, through the graph
    for solo in (1, (1 << 64) - 1):
        b.set_tuning(solo_paths=solo, max_paths=1 << 20, dda_budget=16, refill_min=16, refill_chunk=128)
        for c in ctxs:
            c.reset_accum()
        b.render(seeds, counts, graph=True)
        for i, c in enumerate(ctxs):
            assert np.array_equal(c.read_accum().view(np.uint32), ref[i].view(np.uint32)), f"solo {solo} scene {i}"
    with pytest.raises(api.VTError):
        b.set_tuning(max_paths=100)
    with pytest.raises(api.VTError):
        b.set_tuning(refill_chunk=10)
    b.close()
    for c in ctxs:
        c.close()


def test_context_destroyed_before_its_batch():
    from voxeltoy_amd import api, workloads
    wls = mixed_scenes()[:2]
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    b = api.Batch(ctxs)
    ctxs[1].close()
    with pytest.raises(api.VTError):
        b.render([1, 2], [1, 1])
    b.close()
    ctxs[0].render(wls[0].seed, 0, 1)
    assert ctxs[0].read_accum()[..., 3].min() == 1.0
    ctxs[0].close()


def test_batch_workload_at_full_size_equals_single_renders():
    """The benchmarked batch at its full size (64 scenes, grids 32^3-512^3,
    images 128^2-1024^2, 1-64 spp, 4 bounces; 257 M paths): the HIP-graph
    batch gives every scene exactly its own single-scene render."""
    from voxeltoy_amd import api, workloads
    wls = workloads.batch(64)
    assert len(wls) == 64
    ctxs = []
    for wl in wls:
        c = api.Context(0)
        workloads.load(c, wl)
        ctxs.append(c)
    seeds = [wl.seed for wl in wls]
    counts = [wl.spp for wl in wls]
    b = api.Batch(ctxs)
    b.render(seeds, counts, graph=True, stats=True)
    assert b.stats()["samples"] == sum(wl.width * wl.height * wl.spp for wl in wls)
    batched = [c.read_accum() for c in ctxs]
    for i, (c, wl) in enumerate(zip(ctxs, wls)):
        c.reset_accum()
        c.render(wl.seed, 0, wl.spp)
        assert np.array_equal(c.read_accum().view(np.uint32), batched[i].view(np.uint32)), f"scene {i} differs"
    b.close()
    for c in ctxs:
        c.close()
