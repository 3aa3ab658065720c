"""ctypes binding of include/vt_renderer.h: the renderer object (a mirror of
the reference's Renderer class, src/renderer/renderer.h:27-122).

A Renderer made with device=-1 has no GPU context: volume loading, camera
and settings work, render() raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api

RR_SAMPLES_PENDING, RR_FINISHED_RENDERING = 0, 1
PA_SELECT_FOCAL_POINT, PA_SELECT_ACTIVE_VOXEL, PA_ADD_VOXEL, PA_REMOVE_VOXEL = 0, 1, 2, 3
CLM_PINHOLE, CLM_THIN_LENS, CLM_ORTHOGRAPHIC = 0, 1, 2


class Settings(C.Structure):
    _fields_ = [("max_bounces", C.c_int32), ("max_samples", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("viewport", C.c_int32 * 4), ("wireframe_opacity", C.c_float),
                ("wireframe_thickness", C.c_float), ("background_image", C.c_char_p),
                ("background_top", C.c_float * 3), ("background_bottom", C.c_float * 3),
                ("background_rotation_degrees", C.c_int32), ("exposure", C.c_float), ("gamma", C.c_float)]


class CameraState(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("fov_y", C.c_float), ("near_d", C.c_float),
                ("far_d", C.c_float), ("focal_length", C.c_float), ("focal_distance", C.c_float),
                ("lens_radius", C.c_float), ("film", C.c_float * 2), ("lens_model", C.c_int32)]

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = np.array(list(v), np.float32) if hasattr(v, "__len__") else v
        return d


_R = C.c_void_p
_f3 = C.POINTER(C.c_float)
SIGNATURES = {
    "vtr_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "vtr_destroy": (None, [_R]),
    "vtr_get_status": (C.c_char_p, [_R]),
    "vtr_context": (C.c_void_p, [_R]),
    "vtr_initialize": (C.c_int, [_R, C.c_char_p]),
    "vtr_resize_frame": (C.c_int, [_R] + [C.c_int32] * 6),
    "vtr_render": (C.c_int, [_R]),
    "vtr_reset_render": (C.c_int, [_R]),
    "vtr_set_seed": (C.c_int, [_R, C.c_uint64]),
    "vtr_number_samples": (C.c_int32, [_R]),
    "vtr_load_vox_file": (C.c_int, [_R, C.c_char_p]),
    "vtr_load_mesh": (C.c_int, [_R, C.c_char_p]),
    "vtr_load_volume": (C.c_int, [_R, C.c_int32, C.c_int32, C.c_int32, api.i32p, C.c_void_p, C.c_int32,
                                  C.c_void_p, C.c_int32]),
    "vtr_get_volume": (C.c_int, [_R, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vtr_get_emissive": (C.c_int, [_R, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "vtr_save_image": (C.c_int, [_R, C.c_char_p]),
    "vtr_save_checkpoint": (C.c_int, [_R, C.c_char_p]),
    "vtr_load_checkpoint": (C.c_int, [_R, C.c_char_p]),
    "vtr_get_render_settings": (C.c_int, [_R, C.POINTER(Settings)]),
    "vtr_update_render_settings": (C.c_int, [_R, C.POINTER(Settings)]),
    "vtr_camera_get": (C.c_int, [_R, C.POINTER(CameraState)]),
    "vtr_camera_set_eye_target": (C.c_int, [_R, api.f32p, api.f32p]),
    "vtr_camera_look_at": (C.c_int, [_R, api.f32p]),
    "vtr_camera_set_distance_from_target": (C.c_int, [_R, C.c_float]),
    "vtr_camera_orbit": (C.c_int, [_R, C.c_float, C.c_float]),
    "vtr_camera_focus_on_bounds": (C.c_int, [_R]),
    "vtr_camera_set_lens_model": (C.c_int, [_R, C.c_int32]),
    "vtr_camera_set_focal_length": (C.c_int, [_R, C.c_float]),
    "vtr_camera_set_film_size": (C.c_int, [_R, C.c_float, C.c_float]),
    "vtr_camera_set_lens_radius": (C.c_int, [_R, C.c_float]),
    "vtr_camera_set_fstop": (C.c_int, [_R, C.c_float]),
    "vtr_camera_set_fov_y": (C.c_int, [_R, C.c_float]),
    "vtr_camera_set_controller": (C.c_int, [_R, C.c_int32]),
    "vtr_camera_block": (C.c_int, [_R, C.POINTER(api.Camera)]),
    "vtr_on_mouse_move": (C.c_int, [_R, C.c_int32, C.c_int32, C.c_int32]),
    "vtr_on_key_press": (C.c_int, [_R, C.c_int32]),
    "vtr_request_action": (C.c_int, [_R, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32]),
    "vtr_focal_distance": (C.c_float, [_R]),
    "vtr_selected_voxel": (C.c_int, [_R, api.i32p]),
    "vtr_cpu_voxelize": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32,
                                   C.c_void_p]),
    "vtr_load_obj": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p, C.c_int64,
                               C.POINTER(C.c_int64), C.c_void_p, C.c_void_p]),
    "vtr_get_material_data": (C.c_int, [_R, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "vtr_update_material_color": (C.c_int, [_R, C.c_uint32, api.f32p]),
    "vtr_update_material_value": (C.c_int, [_R, C.c_uint32, C.c_float]),
}

_bound = False


def lib() -> C.CDLL:
    global _bound
    L = api.lib()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


class RendererError(api.VTError):
    pass


def load_obj(path: str):
    """MeshLoader::loadFromOBJ: (vertices (n, 3) float32, triangles (m, 3)
    uint32, bmin, bmax)."""
    nv, nt = C.c_int64(), C.c_int64()
    if lib().vtr_load_obj(str(path).encode(), None, 0, C.byref(nv), None, 0, C.byref(nt), None, None) != 0:
        raise RendererError(f"cannot load {path}")
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    if lib().vtr_load_obj(str(path).encode(), v.ctypes.data, nv.value, C.byref(nv), t.ctypes.data, nt.value,
                          C.byref(nt), bmin.ctypes.data, bmax.ctypes.data) != 0:
        raise RendererError(f"cannot load {path}")
    return v, t, bmin, bmax


def cpu_voxelize(verts, tris, M, res, mode: int = 0) -> np.ndarray:
    """CPUVoxelizer on the host threads: occupancy (res z, y, x) with 1 at the
    covered voxels of the mesh placed by the row-major transform M."""
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    m = np.ascontiguousarray(M, np.float32).reshape(-1)
    r = np.asarray(res, np.int32)
    occ = np.zeros((int(r[2]), int(r[1]), int(r[0])), np.uint8)
    if lib().vtr_cpu_voxelize(v.ctypes.data, len(v), t.ctypes.data, len(t), m.ctypes.data, r.ctypes.data, mode,
                              occ.ctypes.data) != 0:
        raise RendererError("cpu_voxelize: bad arguments")
    return occ


class Renderer:
    """voxeltoy::Renderer through the C ABI; method names follow the
    reference (snake_case)."""

    def __init__(self, device: int = 0):
        self._r = C.c_void_p()
        if lib().vtr_create(device, C.byref(self._r)) != 0:
            raise RendererError("vtr_create failed")

    def _check(self, code: int):
        if code < 0:
            raise RendererError(f"renderer error {code}: {self.status()}")
        return code

    def close(self):
        if self._r:
            lib().vtr_destroy(self._r)
            self._r = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def status(self) -> str:
        return lib().vtr_get_status(self._r).decode()

    def context_handle(self):
        return lib().vtr_context(self._r)

    def initialize(self, shader_path: str = ""):
        self._check(lib().vtr_initialize(self._r, shader_path.encode()))

    def resize_frame(self, w, h, vx=0, vy=0, vw=None, vh=None):
        self._check(lib().vtr_resize_frame(self._r, w, h, vx, vy, w if vw is None else vw, h if vh is None else vh))

    def render(self) -> int:
        return self._check(lib().vtr_render(self._r))

    def reset_render(self):
        self._check(lib().vtr_reset_render(self._r))

    def set_seed(self, seed: int):
        self._check(lib().vtr_set_seed(self._r, seed))

    def number_samples(self) -> int:
        return lib().vtr_number_samples(self._r)

    def load_vox_file(self, path: str):
        self._check(lib().vtr_load_vox_file(self._r, str(path).encode()))

    def load_mesh(self, path: str):
        self._check(lib().vtr_load_mesh(self._r, str(path).encode()))

    def load_volume(self, offsets: np.ndarray, mat: np.ndarray, emissive=None):
        o = np.ascontiguousarray(offsets, np.int32)
        nz, ny, nx = o.shape
        m = np.ascontiguousarray(mat, np.float32)
        e = np.ascontiguousarray([] if emissive is None else emissive, np.int32)
        self._check(lib().vtr_load_volume(self._r, nx, ny, nz, o.reshape(-1), m.ctypes.data_as(C.c_void_p), m.size,
                                          e.ctypes.data_as(C.c_void_p), e.size))

    def volume(self):
        """(offsets[z, y, x], bmin, bmax) of the host copy."""
        res = np.zeros(3, np.int32)
        bmin = np.zeros(3, np.float32)
        bmax = np.zeros(3, np.float32)
        self._check(lib().vtr_get_volume(self._r, res.ctypes.data, bmin.ctypes.data, bmax.ctypes.data, None))
        off = np.zeros((res[2], res[1], res[0]), np.int32)
        self._check(lib().vtr_get_volume(self._r, None, None, None, off.ctypes.data))
        return off, bmin, bmax

    def emissive(self) -> np.ndarray:
        n = C.c_int32()
        self._check(lib().vtr_get_emissive(self._r, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.int32)
        self._check(lib().vtr_get_emissive(self._r, out.ctypes.data, n.value, C.byref(n)))
        return out

    def material_data(self) -> np.ndarray:
        n = C.c_int32()
        self._check(lib().vtr_get_material_data(self._r, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        self._check(lib().vtr_get_material_data(self._r, out.ctypes.data, n.value, C.byref(n)))
        return out

    def save_image(self, path: str):
        self._check(lib().vtr_save_image(self._r, str(path).encode()))

    def save_checkpoint(self, path: str):
        self._check(lib().vtr_save_checkpoint(self._r, str(path).encode()))

    def load_checkpoint(self, path: str):
        self._check(lib().vtr_load_checkpoint(self._r, str(path).encode()))

    def render_settings(self) -> Settings:
        s = Settings()
        self._check(lib().vtr_get_render_settings(self._r, C.byref(s)))
        return s

    def update_render_settings(self, s: Settings):
        self._check(lib().vtr_update_render_settings(self._r, C.byref(s)))

    def camera(self) -> dict:
        c = CameraState()
        self._check(lib().vtr_camera_get(self._r, C.byref(c)))
        return c.as_dict()

    def camera_block(self) -> api.Camera:
        c = api.Camera()
        self._check(lib().vtr_camera_block(self._r, C.byref(c)))
        return c

    def set_eye_target(self, eye, target):
        self._check(lib().vtr_camera_set_eye_target(self._r, np.asarray(eye, np.float32),
                                                     np.asarray(target, np.float32)))

    def look_at(self, target):
        self._check(lib().vtr_camera_look_at(self._r, np.asarray(target, np.float32)))

    def set_distance_from_target(self, d: float):
        self._check(lib().vtr_camera_set_distance_from_target(self._r, d))

    def orbit(self, theta: float, phi: float):
        self._check(lib().vtr_camera_orbit(self._r, theta, phi))

    def focus_on_bounds(self):
        self._check(lib().vtr_camera_focus_on_bounds(self._r))

    def set_lens_model(self, m: int):
        self._check(lib().vtr_camera_set_lens_model(self._r, m))

    def set_focal_length(self, f: float):
        self._check(lib().vtr_camera_set_focal_length(self._r, f))

    def set_film_size(self, w: float, h: float):
        self._check(lib().vtr_camera_set_film_size(self._r, w, h))

    def set_lens_radius(self, r: float):
        self._check(lib().vtr_camera_set_lens_radius(self._r, r))

    def set_fstop(self, f: float):
        self._check(lib().vtr_camera_set_fstop(self._r, f))

    def set_fov_y(self, f: float):
        self._check(lib().vtr_camera_set_fov_y(self._r, f))

    def set_camera_controller(self, mode: int):
        """0 = orbit (CCM_ORBIT), 1 = fly (CCM_FLY)."""
        self._check(lib().vtr_camera_set_controller(self._r, mode))

    def on_mouse_move(self, dx: int, dy: int, buttons: int) -> bool:
        return bool(self._check(lib().vtr_on_mouse_move(self._r, dx, dy, buttons)))

    def on_key_press(self, key: int) -> bool:
        return bool(self._check(lib().vtr_on_key_press(self._r, key)))

    def request_action(self, x, y, dx=0.0, dy=0.0, action=PA_SELECT_FOCAL_POINT, restart=True):
        self._check(lib().vtr_request_action(self._r, x, y, dx, dy, action, int(restart)))

    def focal_distance(self) -> float:
        return float(lib().vtr_focal_distance(self._r))

    def selected_voxel(self) -> np.ndarray:
        v = np.zeros(3, np.int32)
        self._check(lib().vtr_selected_voxel(self._r, v))
        return v

    def update_material_color(self, offset: int, rgb):
        self._check(lib().vtr_update_material_color(self._r, offset, np.asarray(rgb, np.float32)))

    def update_material_value(self, offset: int, v: float):
        self._check(lib().vtr_update_material_value(self._r, offset, v))

    def read_accum(self) -> np.ndarray:
        """The running mean of the renderer's context (H, W, 4), row 0 = bottom."""
        s = self.render_settings()
        out = np.zeros((s.height, s.width, 4), np.float32)
        code = api.lib().vt_read_accum(self.context_handle(), out)
        if code != 0:
            raise RendererError(api.lib().vt_last_error().decode())
        return out
