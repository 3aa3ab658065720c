"""The allocation-free picking entry point (vt_trace_rays_noalloc) against
vt_trace_rays, through the C ABI: bit-identical records on two calls in a
row, the second with fewer rays, and the staging limit enforced on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_RAYS = 16384          # VT_NOALLOC_MAX_RAYS in include/vt_api.h


def _rays(ctx, n, seed):
    rng = np.random.default_rng(seed)
    bmin, bmax, _ = ctx.bounds()
    o = (bmin + rng.uniform(-0.2, 1.2, (n, 3)) * (bmax - bmin)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d[: n // 8, 1:] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def _call(vtapi, fn, ctx, o, d):
    n = len(o)
    out = np.full((n, 8), -7, np.int32)
    rc = getattr(vtapi.lib(), fn)(ctx._c, n, np.ascontiguousarray(o).reshape(-1),
                                  np.ascontiguousarray(d).reshape(-1), out.reshape(-1))
    return rc, out


def test_noalloc_trace_rays_equals_allocating_entry_point(ctx, vtapi, orc):
    from voxeltoy_amd import workloads
    wl = workloads.small(2)
    workloads.load(ctx, wl)
    for n, seed in ((MAX_RAYS, 5), (1237, 6)):
        o, d = _rays(ctx, n, seed)
        print(f"noalloc: {n} rays, seed {seed}", flush=True)
        rc_a, ref = _call(vtapi, "vt_trace_rays", ctx, o, d)
        rc_b, got = _call(vtapi, "vt_trace_rays_noalloc", ctx, o, d)
        assert rc_a == 0 and rc_b == 0, vtapi.lib().vt_last_error()
        bad = np.any(got != ref, axis=-1)
        assert not bad.any(), f"{bad.sum()} of {n} records differ"
        assert (ref[:, 0] == 2).sum() > n // 20          # the rays do hit voxels
    # the second, smaller call also matches the oracle
    sc = wl.scene
    o_sc = orc.OracleScene(sc.offsets(), sc.mat, sc.emissive)
    o_sc.set_camera(**wl.camera_arrays())
    assert np.array_equal(got, o_sc.trace_rays(o, d))


def test_noalloc_trace_rays_limits(ctx, vtapi):
    from voxeltoy_amd import workloads
    wl = workloads.small(0)
    workloads.load(ctx, wl)
    o, d = _rays(ctx, MAX_RAYS + 1, 9)
    rc, out = _call(vtapi, "vt_trace_rays_noalloc", ctx, o, d)
    assert rc == -4                                        # VT_ERR_LIMIT, nothing launched
    assert b"staging limit" in vtapi.lib().vt_last_error()
    assert np.all(out == -7)
    rc, out = _call(vtapi, "vt_trace_rays_noalloc", ctx, o[:0], d[:0])
    assert rc == 0
    # still usable after the rejected call
    rc, got = _call(vtapi, "vt_trace_rays_noalloc", ctx, o[:64], d[:64])
    _, ref = _call(vtapi, "vt_trace_rays", ctx, o[:64], d[:64])
    assert rc == 0 and np.array_equal(got, ref)
