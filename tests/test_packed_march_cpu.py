"""CPU emulation of the packed-coordinate wave march (advance_wave_packed in
voxeltoy_amd/csrc/vt_device.h) for one lane: the packed coordinates
P = (x+4) | (y+4) << 10 | (z+4) << 20, the block key P & 0x3fcff3fc, the
block's linear index bx + bxp (by + byp bz) taken from the key, the block
bit (LDS form) or directly the bit-reversed mask (built as
block_masks_reversed_kernel builds it), the in-block index as the byte sum
of P & 0x00300c03, and the compare/add step.  Every read must fall inside
the padded block arrays, and hit voxel and step count must equal the
oracle's raymarch (src/shaders/shared/dda.h:7-61), also when the march is
resumed every 32 iterations with the non-empty flag carried in bit 31 of
the key."""
import numpy as np
import pytest

F = np.float32
KEY_MASK = 0x3FCFF3FC
IDX_MASK = 0x00300C03


def padded_masks(occ):
    """occp of block_masks_padded_kernel: (kz+2, ky+2, kx+2) uint64 masks,
    border blocks all set, voxels outside the grid set; bit x+4y+16z."""
    nz, ny, nx = occ.shape
    kx, ky, kz = (nx + 3) // 4, (ny + 3) // 4, (nz + 3) // 4
    full = np.ones(((kz + 2) * 4, (ky + 2) * 4, (kx + 2) * 4), bool)
    full[4:4 + nz, 4:4 + ny, 4:4 + nx] = occ
    b = full.reshape(kz + 2, 4, ky + 2, 4, kx + 2, 4).transpose(0, 2, 4, 1, 3, 5)   # bz,by,bx,lz,ly,lx
    bits = b.reshape(kz + 2, ky + 2, kx + 2, 64).astype(np.uint64)
    weights = np.left_shift(np.uint64(1), np.arange(64, dtype=np.uint64))
    return (bits * weights).sum(axis=-1, dtype=np.uint64), (kx, ky, kz)


def reversed_masks(occp):
    """block_masks_reversed_kernel: each padded mask with its bits reversed,
    in the linear padded block order."""
    vals = occp.reshape(-1)
    rev = np.zeros_like(vals)
    for b in range(64):
        rev |= ((vals >> np.uint64(b)) & np.uint64(1)) << np.uint64(63 - b)
    return rev


def march(occr, k, res, bmin, bmax, o, d, max_steps, budget, lds):
    """One lane of advance_wave_packed, called repeatedly as the traversal
    kernel does.  Returns (hit, (x, y, z), steps, max index read)."""
    res_f = [F(r) for r in res]
    ext = [F(1.0) / (F(bmax[i]) - F(bmin[i])) for i in range(3)]
    sgn_d = [F(np.sign(F(v))) for v in d]
    oo = [F(o[i]) + sgn_d[i] * F(0.001) for i in range(3)]
    vo = [((oo[i] - F(bmin[i])) * ext[i]) * res_f[i] for i in range(3)]
    vp = [F(np.floor(v)) for v in vo]
    if any(v < 0 or v >= r for v, r in zip(vp, res_f)):
        return None
    dc = [F(1e-5) if F(1e-5) >= abs(F(v)) else F(v) for v in d]
    inc = [F(1.0) / v for v in dc]
    sg = [int(np.sign(v)) for v in dc]
    dis = [(((vp[i] - vo[i]) + F(0.5)) + F(sg[i]) * F(0.5)) * inc[i] for i in range(3)]
    ax = [abs(v) for v in inc]
    x, y, z = (int(v) for v in vp)
    bxp, byp = k[0] + 2, k[1] + 2
    steps, key, mask, max_idx = 0, 0xFFFFFFFF, 0, -1
    inc_p = [sg[0], sg[1] * 1024, sg[2] * (1 << 20)]
    while True:
        room = max_steps - steps
        n = max(0, min(budget, room))
        p = (x + 4) | ((y + 4) << 10) | ((z + 4) << 20)
        ne = key >= 0x80000000
        key &= 0x7FFFFFFF
        it, hit = 0, False
        while it < n:
            t0 = p & KEY_MASK
            if t0 != key:
                key = t0
                bx, by, bz = (t0 >> 2) & 255, (t0 >> 12) & 255, t0 >> 22
                i = bx + bxp * (by + byp * bz)
                assert 0 <= i < len(occr), f"block read {i} outside {len(occr)} entries"
                max_idx = max(max_idx, i)
                if lds and occr[i] == 0:          # the block's bit: the mask is not loaded
                    ne = False
                else:
                    mask = int(occr[i])
                    ne = mask != 0
            if ne:
                t = p & IDX_MASK
                s = (t & 255) + ((t >> 8) & 255) + ((t >> 16) & 255) + (t >> 24)
                if (mask << s) & (1 << 63):
                    hit = True
                    break
            m = min(dis)
            mx, my, mz = dis[0] == m, dis[1] == m, dis[2] == m
            if mx:
                dis[0] = dis[0] + ax[0]
                p += inc_p[0]
            if my:
                dis[1] = dis[1] + ax[1]
                p += inc_p[1]
            if mz:
                dis[2] = dis[2] + ax[2]
                p += inc_p[2]
            it += 1
        steps += it
        x, y, z = (p & 1023) - 4, ((p >> 10) & 1023) - 4, (p >> 20) - 4
        key |= 0x80000000 if ne else 0
        inside = 0 <= x < res[0] and 0 <= y < res[1] and 0 <= z < res[2]
        if hit:
            return inside, (x, y, z), steps, max_idx
        if steps >= max_steps:
            return False, (x, y, z), steps, max_idx


def check_grid(orc, occ, n_rays, seed, budget=32, axis_rays=True):
    nz, ny, nx = occ.shape
    offsets = np.where(occ, 0, -1).astype(np.int32)
    mat = np.array([0, 0, 0, 0, 0.5, 0.5, 0.5, 0], np.float32)
    sc = orc.OracleScene(offsets, mat)
    occp, k = padded_masks(occ)
    occr = reversed_masks(occp)
    res = (nx, ny, nz)
    max_steps = orc.lib().orc_max_steps(np.array(res, np.int32))
    rng = np.random.default_rng(seed)
    lo, hi = sc.bmin.astype(np.float64), sc.bmax.astype(np.float64)
    checked = 0
    for r in range(n_rays):
        o = (lo + rng.uniform(0.0, 1.0, 3) * (hi - lo)).astype(np.float32)
        d = rng.standard_normal(3)
        if axis_rays and r % 5 == 0:
            d = np.zeros(3)
            d[r % 3] = 1.0 if r % 2 else -1.0
        d = (d / np.linalg.norm(d)).astype(np.float32)
        got = march(occr, k, res, sc.bmin, sc.bmax, o, d, max_steps, budget, lds=bool(r % 2))
        vp = np.zeros(3, np.int32)
        st = np.zeros(1, np.int32)
        wrote = np.zeros(1, np.int32)
        import ctypes as C
        hit = orc.lib().orc_raymarch(C.byref(sc.s), o, d, max_steps, vp, st.ctypes.data_as(C.POINTER(C.c_int32)),
                                     wrote.ctypes.data_as(C.POINTER(C.c_int32)))
        if got is None:
            assert wrote[0] == 0
            continue
        ghit, gvp, gsteps, _ = got
        assert bool(ghit) == bool(hit), (r, o, d)
        assert gsteps == st[0], (r, gsteps, st[0])
        assert tuple(gvp) == tuple(vp), (r, gvp, vp)
        checked += 1
    return checked


@pytest.mark.parametrize("shape,seed", [((16, 16, 16), 1), ((13, 18, 21), 2), ((24, 9, 30), 3)])
def test_packed_march_matches_oracle_and_stays_in_bounds(orc, shape, seed):
    rng = np.random.default_rng(seed)
    occ = rng.random(shape) < 0.08
    assert check_grid(orc, occ, 120, seed) > 60


def test_packed_march_at_the_largest_packed_extent(orc):
    """1016 voxels along x (the largest extent the packed march takes):
    axis rays reach x = -1 and x = 1016, padded blocks 0 and 255."""
    occ = np.zeros((5, 7, 1016), bool)
    occ[2, 3, 500] = True
    assert check_grid(orc, occ, 30, 7, budget=64) > 15


def test_key_and_index_positions():
    """The key keeps each field's block coordinate at bits 2, 12 and 22 (the
    fields the reload extracts), and the byte sum gives the in-block bit."""
    for (x, y, z) in [(0, 0, 0), (3, 1022, 5), (1019, 7, 1019), (255 * 4, 255 * 4, 255 * 4)]:
        p = x | (y << 10) | (z << 20)
        k = p & KEY_MASK
        assert ((k >> 2) & 255, (k >> 12) & 255, k >> 22) == (x >> 2, y >> 2, z >> 2)
        t = p & IDX_MASK
        assert (t & 255) + ((t >> 8) & 255) + ((t >> 16) & 255) == (x & 3) + 4 * (y & 3) + 16 * (z & 3)
