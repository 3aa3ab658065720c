"""Named render workloads and a helper that loads one into a Context.

  small  : 64^3 sphere + heightfield, all Lambert, 256x256 pinhole camera at 8
           orbit angles, constant-white environment, 4 spp, 4 bounces.
  full   : 256^3 terrain + boxes, 70/20/10 Lambert/metal/emissive, 1024x512
           FP32 sky+sun environment (importance sampled), 1024x1024 thin-lens
           camera looking down onto the terrain (lens radius 0.02 x grid
           extent, random focus depth), 16 spp, 8 bounces.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import api, scenes


@dataclass
class Workload:
    name: str
    scene: scenes.Scene
    cam: scenes.CameraSetup
    width: int
    height: int
    spp: int
    bounces: int
    seed: int = 1234
    env_image: np.ndarray | None = None
    env_top: tuple = (1.0, 1.0, 1.0)
    env_bottom: tuple = (1.0, 1.0, 1.0)
    env_rotation: float = 0.0
    cdf: tuple | None = None           # (cdf_u, cdf_v, integral)
    bmin: np.ndarray = field(default=None)
    bmax: np.ndarray = field(default=None)

    def camera_arrays(self):
        """Matrices as Renderer::updateCamera builds them, plus viewport and
        window depth of the full-screen quad."""
        c = self.cam
        mv, inv_mv, proj, inv_proj = api.camera_matrices(c.eye, c.target, c.fov_y, c.near, c.far, c.lens_model,
                                                         self.width, self.height)
        vp = np.array([0, 0, self.width, self.height], np.float32)
        return dict(inv_mv=inv_mv, proj=proj, inv_proj=inv_proj, viewport=vp, win_z=scenes.window_z(c.near),
                    lens_model=c.lens_model, lens_radius=c.lens_radius, focal_distance=c.focal_distance)


def reference_bounds(shape):
    nz, ny, nx = shape
    vs = np.float32(1000.0) / np.float32(max(nx, ny, nz))
    size = np.array([vs * np.float32(nx), vs * np.float32(ny), vs * np.float32(nz)], np.float32)
    return -size * np.float32(0.5), size * np.float32(0.5)


SMALL_THETA = math.radians(120.0)   # eye 30 degrees above the horizon


def small(angle: int = 0, width: int = 256, height: int = 256, seed_scene: int = 1, n: int = 64) -> Workload:
    sc = scenes.small_scene(seed_scene, n=n)
    bmin, bmax = reference_bounds(sc.shape)
    cam = scenes.orbit_camera(bmin, bmax, SMALL_THETA, math.radians(45.0 * angle), width, height)
    return Workload("small", sc, cam, width, height, spp=4, bounces=4, bmin=bmin, bmax=bmax)


def full(width: int = 1024, height: int = 1024, seed_scene: int = 2, n: int = 256, env_w: int = 1024,
         env_h: int = 512, spp: int = 16, bounces: int = 8) -> Workload:
    sc = scenes.full_scene(seed_scene, n=n)
    offsets = sc.offsets()
    em = api.prune_emissive(offsets, sc.emissive)
    sc.emissive = em
    bmin, bmax = reference_bounds(sc.shape)
    rng = np.random.default_rng(seed_scene + 100)
    extent = float(np.max(bmax - bmin))
    # looking down 40 degrees onto the terrain from half the focusOnBounds distance
    cam = scenes.orbit_camera(bmin, bmax, math.radians(130.0), math.radians(30.0), width, height, lens_model=1,
                              lens_radius=0.02 * extent)
    target = cam.target.copy()
    target[1] = bmin[1] + 0.2 * extent
    cam.eye = (target + (cam.eye - cam.target) * 0.5).astype(np.float32)
    cam.target = target.astype(np.float32)
    dist = float(np.linalg.norm(cam.target - cam.eye))
    cam.focal_distance = float(rng.uniform(dist - 0.3 * extent, dist + 0.2 * extent))
    img = scenes.sky_environment(seed_scene + 1, env_w, env_h)
    cdf = api.env_build_cdf(img)
    return Workload("full", sc, cam, width, height, spp=spp, bounces=bounces, env_image=img, cdf=cdf, bmin=bmin,
                    bmax=bmax)


def batch(count: int = 64, seed: int = 7, bounces: int = 4, n_range=(32, 512), res_range=(128, 1024),
          spp_range=(1, 64), occ_range=(0.05, 0.6)) -> list:
    """Many independent scenes: grid edge, image edge and spp drawn
    log-uniformly from their ranges, occupancy uniformly; materials as in
    `full`; gradient environment; pinhole camera at the focusOnBounds
    distance from a random direction above the horizon."""
    rng = np.random.default_rng(seed)

    def log_uniform(lo, hi):
        return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))

    out = []
    for i in range(count):
        n = log_uniform(*n_range)
        res = log_uniform(*res_range)
        spp = log_uniform(*spp_range)
        occ = float(rng.uniform(*occ_range))
        sc = scenes.mixed_scene(rng, n, occ)
        offsets = sc.offsets()
        sc.emissive = api.prune_emissive(offsets, sc.emissive)
        bmin, bmax = reference_bounds(sc.shape)
        cam = scenes.orbit_camera(bmin, bmax, math.radians(rng.uniform(100, 150)), rng.uniform(0, 2 * math.pi),
                                  res, res)
        out.append(Workload(f"batch{i}", sc, cam, res, res, spp=spp, bounces=bounces, seed=1000 + i, bmin=bmin,
                            bmax=bmax, env_top=(1.6, 1.7, 1.9), env_bottom=(0.35, 0.3, 0.25)))
    return out


def load(ctx: api.Context, wl: Workload, fp64: bool = False, int32_grid: bool = False):
    """Upload scene, camera, environment and settings of `wl` into ctx."""
    sc = wl.scene
    if int32_grid:
        ctx.set_volume(sc.offsets(), sc.mat, sc.emissive)
    else:
        ctx.set_volume_u8(sc.ids, sc.id_offset, sc.mat, sc.emissive)
    ctx.set_camera(**wl.camera_arrays())
    if wl.env_image is not None:
        cu, cv, integ = wl.cdf
        ctx.set_environment(wl.env_top, wl.env_bottom, wl.env_image, cu, cv, integ, wl.env_rotation)
    else:
        ctx.set_environment(wl.env_top, wl.env_bottom, rotation=wl.env_rotation)
    ctx.set_settings(max_bounces=wl.bounces)
    ctx.set_frame(wl.width, wl.height, fp64=fp64)
