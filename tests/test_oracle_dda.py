"""Pins the oracle's traversal (dda.h raymarch/traverse, aabb.h, the basis of
coordinates.h) by closed forms, invariants and an exact-rational DDA."""
from fractions import Fraction

import numpy as np
import pytest

f32 = np.float32


def make_scene(orc, n=16, occupied=(), shape=None):
    nz, ny, nx = shape or (n, n, n)
    off = np.full((nz, ny, nx), -1, np.int32)
    for (x, y, z) in occupied:
        off[z, y, x] = 0
    mat = np.array([0, 0, 0, 0, 0.5, 0.5, 0.5], np.float32)
    return orc.OracleScene(off, mat), off


def march(orc, sc, o, d):
    vp = np.zeros(3, np.int32)
    steps, wrote = orc.C.c_int32(), orc.C.c_int32()
    ms = orc.lib().orc_max_steps(np.array(sc.offsets.shape[::-1], np.int32))
    hit = orc.lib().orc_raymarch(orc.C.byref(sc.s), f32(o), f32(d), ms, vp, orc.C.byref(steps), orc.C.byref(wrote))
    return hit, tuple(vp), steps.value, wrote.value


def traverse(orc, sc, o, d):
    vp = np.zeros(3, np.int32)
    steps, kind = orc.C.c_int32(), orc.C.c_int32()
    orc.lib().orc_traverse(orc.C.byref(sc.s), f32(o), f32(d), vp, orc.C.byref(steps), orc.C.byref(kind))
    return kind.value, tuple(vp), steps.value


@pytest.fixture(scope="module")
def C(orc):
    import ctypes
    orc.C = ctypes
    return ctypes


def test_max_steps(orc, C):
    ms = lambda r: orc.lib().orc_max_steps(np.array(r, np.int32))
    assert ms([64, 64, 64]) == 222
    assert ms([256, 256, 256]) == 888
    assert ms([512, 512, 512]) == 1774
    assert ms([126, 20, 126]) == 2 * int(np.ceil(np.sqrt(126 ** 2 * 2 + 20 ** 2)))


def voxel_center(sc, x, y, z):
    return sc.bmin + (np.array([x, y, z], np.float32) + f32(0.5)) * sc.vsize


@pytest.mark.parametrize("axis,sign", [(0, 1), (0, -1), (1, 1), (2, 1), (2, -1)])
def test_axis_rays_closed_form(orc, C, axis, sign):
    n = 16
    start = [5, 6, 7]
    d = np.zeros(3, np.float32)
    d[axis] = sign
    sc, _ = make_scene(orc, n)
    hit, vp, steps, wrote = march(orc, sc, voxel_center(sc, *start), d)
    # no voxel occupied: the march visits every voxel to the boundary
    expect = (n - start[axis]) if sign > 0 else (start[axis] + 1)
    assert (hit, wrote, steps) == (0, 1, expect)
    for k in (0, 1, 3):
        tgt = list(start)
        tgt[axis] += sign * k
        sc, _ = make_scene(orc, n, occupied=[tuple(tgt)])
        hit, vp, steps, _ = march(orc, sc, voxel_center(sc, *start), d)
        assert hit == 1 and vp == tuple(tgt) and steps == k


def exact_sequence(vo, d, res, max_len=200):
    """Voxels visited by an exact-arithmetic DDA from voxel-space origin vo
    along d; None when two axes cross within a near-tie."""
    vp = [int(np.floor(v)) for v in vo]
    t = []
    inc = []
    step = []
    for i in range(3):
        di = Fraction(float(d[i]))
        s = 1 if di > 0 else -1
        nxt = vp[i] + (1 if s > 0 else 0)
        t.append((Fraction(nxt) - Fraction(float(vo[i]))) / di)
        inc.append(abs(1 / di))
        step.append(s)
    seq = [tuple(vp)]
    while len(seq) < max_len:
        m = min(t)
        close = sorted(t)
        if close[1] - close[0] < Fraction(1, 10 ** 4) * (1 + abs(m)):
            return None
        i = t.index(m)
        vp[i] += step[i]
        t[i] += inc[i]
        seq.append(tuple(vp))
        if not all(0 <= vp[k] < res[k] for k in range(3)):
            break
    return seq


def test_dda_matches_exact_rational_march(orc, C):
    n = 16
    rng = np.random.default_rng(5)
    sc0, _ = make_scene(orc, n)
    res = [n, n, n]
    ext = (f32(1.0) / (sc0.bmax - sc0.bmin)).astype(np.float32)
    checked = 0
    for _ in range(300):
        o = (sc0.bmin + rng.uniform(0.05, 0.95, 3).astype(np.float32) * (sc0.bmax - sc0.bmin)).astype(np.float32)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        if np.min(np.abs(d)) < 0.05:
            continue
        d = d.astype(np.float32)
        o2 = (o + np.sign(d).astype(np.float32) * f32(0.001)).astype(np.float32)
        vo = (((o2 - sc0.bmin).astype(np.float32) * ext).astype(np.float32) * np.float32(n)).astype(np.float32)
        seq = exact_sequence(vo, d, res)
        if seq is None:
            continue
        inside = [v for v in seq if all(0 <= v[k] < n for k in range(3))]
        # empty grid: steps = number of voxels inside, final = first outside
        hit, vp, steps, wrote = march(orc, sc0, o, d)
        assert (hit, wrote) == (0, 1)
        assert steps == len(inside) and vp == seq[len(inside)]
        # adjacency invariant
        for a, b in zip(seq, seq[1:]):
            assert sum(abs(a[k] - b[k]) for k in range(3)) == 1
        for k in sorted(k for k in {0, 1, len(inside) // 2, len(inside) - 1} if k < len(inside)):
            sc, _ = make_scene(orc, n, occupied=[inside[k]])
            hit, vp, steps, _ = march(orc, sc, o, d)
            assert hit == 1 and vp == inside[k] and steps == k
        checked += 1
    assert checked > 150


def test_ground_rule_and_outside_start(orc, C):
    n = 8
    sc, _ = make_scene(orc, n)
    kind, vp, steps = traverse(orc, sc, voxel_center(sc, 3, 2, 4), np.array([0, -1, 0], np.float32))
    assert kind == 3 and vp == (3, -1, 4) and steps == 3
    kind, vp, steps = traverse(orc, sc, voxel_center(sc, 3, 2, 4), np.array([0, 1, 0], np.float32))
    assert kind == 0 and vp == (3, n, 4)
    out = sc.bmax + f32(10)
    hit, vp, steps, wrote = march(orc, sc, out, np.array([1, 1, 1], np.float32))
    assert wrote == 0 and hit == 0
    kind, _, steps = traverse(orc, sc, out, np.array([-1, -1, -1], np.float32))
    assert kind == 0 and steps == -1


def test_aabb_closed_form(orc, C):
    bmin = np.array([-1, -2, -3], np.float32)
    bmax = np.array([1, 2, 3], np.float32)
    A = lambda o, d: orc.lib().orc_aabb(f32(o), f32(d), bmin, bmax)
    assert A([-5, 0, 0], [1, 0, 0]) == pytest.approx(4.0)
    assert A([0, 0, -10], [0, 0, 1]) == pytest.approx(7.0)
    assert A([0, 0, 0], [1, 0, 0]) == 0.0            # inside: clamped to 0
    assert A([5, 0, 0], [1, 0, 0]) == -1.0           # box behind
    assert A([-5, 5, 0], [1, 0, 0]) == -1.0          # passes beside
    assert A([-5, -5, 0], [1, 1, 0]) == pytest.approx(4.0 * np.sqrt(1.0) * 1.0, rel=1e-6)


def test_basis_face_normals(orc, C):
    n = 8
    sc, _ = make_scene(orc, n, occupied=[(4, 4, 4)])
    out = np.zeros(12, np.float32)
    for axis in range(3):
        for sign in (1, -1):
            d = np.zeros(3, np.float32)
            d[axis] = -sign
            o = voxel_center(sc, 4, 4, 4).copy()
            o[axis] += sign * 3 * sc.vsize[axis]
            orc.lib().orc_basis(orc.C.byref(sc.s), np.array([4, 4, 4], np.int32), f32(o), d, out)
            nrm = out[6:9]
            expect = np.zeros(3)
            expect[axis] = sign
            assert np.array_equal(nrm, expect)
            T, B = out[3:6], out[9:12]
            assert abs(np.dot(T, nrm)) < 1e-6 and abs(np.dot(B, nrm)) < 1e-6 and abs(np.dot(T, B)) < 1e-6
            assert np.linalg.norm(T) == pytest.approx(1, abs=1e-6) and np.linalg.norm(B) == pytest.approx(1, abs=1e-6)
            # hit point lies on the face plane
            face = sc.bmin[axis] + (4 + (1 if sign > 0 else 0)) * sc.vsize[axis]
            assert out[axis] == pytest.approx(face, abs=1e-3)
