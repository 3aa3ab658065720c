"""GPU: a wavefront render whose pass halves run on two streams gives the
same accumulation, FP64 sums and ray statistics as the one-stream render
(the second half accumulates after the first)."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_paths", [0, 3_000_000])
def test_two_streams_equal_one_stream(max_paths):
    from voxeltoy_amd import api, scenes, workloads
    sc = scenes.full_scene(7, n=64, n_ids=12, n_boxes=5)
    bmin, bmax = workloads.reference_bounds(sc.shape)
    cam = scenes.orbit_camera(bmin, bmax, math.radians(140), math.radians(35), 1024, 1024)
    wl = workloads.Workload("ov", sc, cam, 1024, 1024, spp=5, bounces=2, seed=33, bmin=bmin, bmax=bmax,
                            env_top=(1.4, 1.5, 1.7), env_bottom=(0.3, 0.3, 0.3))
    out = []
    for streams in (1, 2):
        c = api.Context(0)
        workloads.load(c, wl, fp64=True)
        c.set_render_mode(1, max_paths)
        c.set_overlap(streams)
        c.set_stats(True)
        c.render(wl.seed, 1, wl.spp)       # passes 1..5 (n0 = 1 after a first pass)
        st = c.stats()
        out.append((c.read_accum(), c.read_sum64(), st))
        c.close()
    (a1, s1, t1), (a2, s2, t2) = out
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))
    assert np.array_equal(s1.view(np.uint64), s2.view(np.uint64))
    for k in ("rays", "visits", "interactions", "samples"):
        assert t1[k] == t2[k], k
