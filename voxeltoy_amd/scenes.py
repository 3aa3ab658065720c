"""Seeded synthetic inputs for the render and voxelizer workloads.

Every generator is deterministic in its seed.  Scenes use voxelToy's own
formats: a material-record array (type, emission[3], albedo[3] for
Lambert; type, emission[3], reflectance[3], roughness for metal;
src/renderer/loaders/voxLoader.cpp:195-253) and a uint8 ID grid whose IDs
map to record offsets.  Grids are indexed [z, y, x] (x fastest in memory).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MT_LAMBERT, MT_METAL, MT_PLASTIC = 0, 1, 2

FILM_35MM = 36.0


@dataclass
class Scene:
    ids: np.ndarray                 # (nz, ny, nx) uint8, 0 = empty
    id_offset: np.ndarray           # (256,) int32, -1 = empty
    mat: np.ndarray                 # float32 material records
    emissive: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))

    @property
    def shape(self):
        return self.ids.shape

    def offsets(self) -> np.ndarray:
        """Reference-format int32 material-offset grid (-1 = empty)."""
        return self.id_offset[self.ids.astype(np.int64)].astype(np.int32)


def lambert(emission, albedo):
    return [float(MT_LAMBERT), *map(float, emission), *map(float, albedo)]


def metal(emission, reflectance, roughness):
    return [float(MT_METAL), *map(float, emission), *map(float, reflectance), float(roughness)]


def plastic(emission, diffuse_albedo, roughness):
    """PlasticMaterialData (src/renderer/material/material.h:28-33); the
    shader ignores the diffuse albedo (plastic.h: microfacet, reflectance 1)."""
    return [float(MT_PLASTIC), *map(float, emission), *map(float, diffuse_albedo), float(roughness)]


def build_materials(records):
    """Concatenate records; returns (mat float32, offsets list)."""
    mat, offs = [], []
    for r in records:
        offs.append(len(mat))
        mat.extend(r)
    return np.asarray(mat, np.float32), offs


def value_noise_2d(rng, n, cells):
    """Smooth value noise on an n x n grid with `cells` lattice cells."""
    lat = rng.random((cells + 2, cells + 2))
    t = (np.arange(n) + 0.5) / n * cells
    i = np.floor(t).astype(int)
    f = t - i
    s = f * f * (3 - 2 * f)
    a = lat[np.ix_(i, i)]
    b = lat[np.ix_(i, i + 1)]
    c = lat[np.ix_(i + 1, i)]
    d = lat[np.ix_(i + 1, i + 1)]
    sx = s[None, :]
    sy = s[:, None]
    return (a * (1 - sx) + b * sx) * (1 - sy) + (c * (1 - sx) + d * sx) * sy


def emissive_list(ids, id_offset, mat):
    """Linear indices (x + y*nx + z*nx*ny) of voxels whose material emits."""
    emits = np.zeros(256, bool)
    for i in range(1, 256):
        o = id_offset[i]
        if o >= 0:
            emits[i] = (mat[o + 1] + mat[o + 2] + mat[o + 3]) > 0
    return np.flatnonzero(emits[ids.reshape(-1)]).astype(np.int32)


def small_scene(seed: int = 1, n: int = 64, n_ids: int = 16) -> Scene:
    """64^3 sphere plus value-noise heightfield, ~30% occupied, all Lambert,
    albedo uniform in [0.2, 0.8] per ID."""
    rng = np.random.default_rng(seed)
    h = value_noise_2d(rng, n, 4) * 0.35 * n + 0.08 * n          # heights over (z, x)
    y = np.arange(n)[None, :, None]
    occ = y < h[:, None, :]
    c = np.array([0.5 + 0.1 * rng.standard_normal(), 0.55, 0.5 + 0.1 * rng.standard_normal()]) * n
    r = 0.22 * n
    zz, yy, xx = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    occ |= (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 < r * r
    recs = [lambert((0, 0, 0), rng.uniform(0.2, 0.8, 3)) for _ in range(n_ids)]
    mat, offs = build_materials(recs)
    id_offset = np.full(256, -1, np.int32)
    id_offset[1:n_ids + 1] = offs
    ids = np.where(occ, rng.integers(1, n_ids + 1, size=occ.shape), 0).astype(np.uint8)
    return Scene(ids, id_offset, mat)


def full_scene(seed: int = 2, n: int = 256, n_ids: int = 40, n_boxes: int = 24) -> Scene:
    """Noise terrain plus random boxes; IDs 70% Lambert, 20% metal (roughness
    U[0.05, 0.6], used directly as the Blinn exponent as the reference does),
    10% emissive Lambert (radiance log-normal mu=1, sigma=0.5)."""
    rng = np.random.default_rng(seed)
    h = (0.6 * value_noise_2d(rng, n, 6) + 0.4 * value_noise_2d(rng, n, 17)) * 0.3 * n + 0.05 * n
    y = np.arange(n)[None, :, None]
    occ = y < h[:, None, :]
    for _ in range(n_boxes):
        s = rng.integers(n // 32, n // 8, size=3)
        p = rng.integers(0, n - s)
        p[1] = rng.integers(0, n // 2)
        occ[p[2]:p[2] + s[2], p[1]:p[1] + s[1], p[0]:p[0] + s[0]] = True
    n_lam = int(round(0.7 * n_ids))
    n_met = int(round(0.2 * n_ids))
    n_emi = n_ids - n_lam - n_met
    recs = [lambert((0, 0, 0), rng.uniform(0.2, 0.8, 3)) for _ in range(n_lam)]
    recs += [metal((0, 0, 0), rng.uniform(0.5, 0.95, 3), rng.uniform(0.05, 0.6)) for _ in range(n_met)]
    recs += [lambert(rng.lognormal(1.0, 0.5) * np.ones(3) * rng.uniform(0.7, 1.0, 3), rng.uniform(0.2, 0.8, 3))
             for _ in range(n_emi)]
    mat, offs = build_materials(recs)
    id_offset = np.full(256, -1, np.int32)
    id_offset[1:n_ids + 1] = offs
    ids = np.where(occ, rng.integers(1, n_ids + 1, size=occ.shape), 0).astype(np.uint8)
    sc = Scene(ids, id_offset, mat)
    sc.emissive = emissive_list(ids, id_offset, mat)
    return sc


def mixed_scene(rng, n: int, occupancy: float, n_ids: int = 24) -> Scene:
    """n^3 noise terrain plus boxes filled to roughly `occupancy` of the
    grid; IDs 70% Lambert, 20% metal, 10% emissive as in full_scene."""
    heights = value_noise_2d(rng, n, max(2, n // 32)) * 0.5 + 0.5 * value_noise_2d(rng, n, max(3, n // 12))
    # terrain holds about 80% of the target occupancy, boxes the rest
    h = heights * (1.6 * occupancy * n)
    y = np.arange(n, dtype=np.float32)[None, :, None]
    occ = y < h[:, None, :]
    deficit = occupancy * n ** 3 - float(occ.sum())
    added = 0.0
    while added < deficit and n >= 8:   # box volumes, overlaps ignored
        s = rng.integers(max(1, n // 16), max(2, n // 4), size=3)
        p = rng.integers(0, n - s + 1)
        occ[p[2]:p[2] + s[2], p[1]:p[1] + s[1], p[0]:p[0] + s[0]] = True
        added += float(np.prod(s))
    n_lam = int(round(0.7 * n_ids))
    n_met = int(round(0.2 * n_ids))
    n_emi = n_ids - n_lam - n_met
    recs = [lambert((0, 0, 0), rng.uniform(0.2, 0.8, 3)) for _ in range(n_lam)]
    recs += [metal((0, 0, 0), rng.uniform(0.5, 0.95, 3), rng.uniform(0.05, 0.6)) for _ in range(n_met)]
    recs += [lambert(rng.lognormal(1.0, 0.5) * np.ones(3) * rng.uniform(0.7, 1.0, 3), rng.uniform(0.2, 0.8, 3))
             for _ in range(n_emi)]
    mat, offs = build_materials(recs)
    id_offset = np.full(256, -1, np.int32)
    id_offset[1:n_ids + 1] = offs
    ids = np.zeros(occ.shape, np.uint8)
    ids[occ] = rng.integers(1, n_ids + 1, size=int(occ.sum()), dtype=np.uint8)
    sc = Scene(ids, id_offset, mat)
    sc.emissive = emissive_list(ids, id_offset, mat)
    return sc


def direction_from_uv(u, v, rotation=0.0):
    """coordinates.h:111-116 in float64 (for building environment images)."""
    u, v = np.broadcast_arrays(u, v)
    phi = u * 2 * math.pi - math.pi - rotation
    theta = v * math.pi
    return np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1)


def sky_environment(seed: int = 3, w: int = 1024, h: int = 512, sun_peak: float = 1e4,
                    sun_exponent: float = 4000.0) -> np.ndarray:
    """FP32 equirect (h, w, 3): sky gradient plus a sun lobe of `sun_peak`
    peak radiance.  Row 0 is v = 0 (the +Y pole)."""
    rng = np.random.default_rng(seed)
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(h) + 0.5) / h
    d = direction_from_uv(u[None, :], v[:, None])
    up = np.clip(d[..., 1], -1, 1)
    horizon = np.array([1.0, 0.95, 0.85])
    zenith = np.array([0.25, 0.45, 0.9])
    ground = np.array([0.25, 0.22, 0.2])
    t = np.clip(up, 0, 1)[..., None] ** 0.5
    sky = horizon * (1 - t) + zenith * t
    sky = np.where(up[..., None] >= 0, sky, ground * (1 + up[..., None]) + horizon * (-up[..., None]) * 0.0)
    el = rng.uniform(0.35, 0.9)
    az = rng.uniform(0, 2 * math.pi)
    sun = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    c = np.clip((d * sun).sum(-1), 0, 1)
    lobe = sun_peak * c ** sun_exponent
    img = sky + lobe[..., None] * np.array([1.0, 0.97, 0.9])
    return img.astype(np.float32)


def geodesic_sphere(freq: int):
    """Frequency-`freq` geodesic icosphere: 20*freq^2 triangles, shared
    vertices, unit radius."""
    t = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = {}
    vlist = []

    def key_of(p):
        return tuple(np.round(p / np.linalg.norm(p), 9))

    def vid(p):
        k = key_of(p)
        if k not in verts:
            verts[k] = len(vlist)
            vlist.append(p / np.linalg.norm(p))
        return verts[k]

    tris = []
    for a, b, c in F:
        A, B, Cc = V[a], V[b], V[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                p = A + (B - A) * (i / freq) + (Cc - A) * (j / freq)
                grid[i, j] = vid(p)
        for i in range(freq):
            for j in range(freq - i):
                tris.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j < freq - 1:
                    tris.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    return np.asarray(vlist), np.asarray(tris, np.uint32)


def displaced_icosphere(seed: int = 4, freq: int = 100, amplitude: float = 0.05, radius: float = 1.0):
    """Seeded displaced geodesic icosphere (20*freq^2 triangles; freq=100
    gives 200,000).  Each vertex moves radially by amplitude*radius times a
    smooth noise value in [-1, 1]."""
    rng = np.random.default_rng(seed)
    v, t = geodesic_sphere(freq)
    k = rng.standard_normal((8, 3))
    ph = rng.uniform(0, 2 * math.pi, 8)
    noise = np.zeros(len(v))
    for i in range(8):
        noise += np.sin(v @ (k[i] * (1.5 + i)) + ph[i]) / (1 + i)
    noise /= np.abs(noise).max()
    verts = v * (radius * (1 + amplitude * noise))[:, None]
    return verts.astype(np.float32), t


# ---- camera -------------------------------------------------------------------------------

@dataclass
class CameraSetup:
    eye: np.ndarray
    target: np.ndarray
    fov_y: float
    near: float = 0.1
    far: float = 10000.0
    lens_model: int = 0
    lens_radius: float = 0.0
    focal_distance: float = 99999999.0


def default_fov(width: int, height: int, focal_length_mm: float = 50.0) -> float:
    """CameraParameters: setFocalLength(50) on a 36 mm film; the per-frame
    setFilmSize keeps the vertical field of view
    (src/camera/cameraParameters.cpp:135-163, src/renderer/renderer.cpp:571-579)."""
    f32 = np.float32
    return float(f32(np.arctan2(f32(FILM_35MM * 0.5), f32(focal_length_mm))) * f32(2.0))


def orbit_camera(bmin, bmax, theta: float, phi: float, width: int, height: int, lens_model: int = 0,
                 lens_radius: float = 0.0, focal_distance: float = 99999999.0) -> CameraSetup:
    """Eye on the sphere around the bounds centre at polar angle theta (from
    +Y, orbitAroundTarget) and azimuth phi, at the focusOnBounds distance
    (src/camera/cameraController.cpp:16-23)."""
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    fov = default_fov(width, height)
    center = (bmin + bmax) / 2
    dist = np.linalg.norm(bmax - bmin) / (2 * math.tan(fov / 2))
    sph = np.array([math.sin(theta) * math.cos(phi), math.cos(theta), math.sin(theta) * math.sin(phi)])
    eye = center - dist * sph
    return CameraSetup(eye.astype(np.float32), center.astype(np.float32), fov, lens_model=lens_model,
                       lens_radius=lens_radius, focal_distance=focal_distance)


def window_z(near: float = 0.1) -> float:
    """gl_FragCoord.z of the full-screen quad drawn at z = near under identity
    matrices (src/renderer/renderer.cpp:563-566, 647-656)."""
    f32 = np.float32
    return float(f32(0.5) * f32(near) + f32(0.5))
