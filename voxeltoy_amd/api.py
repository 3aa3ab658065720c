"""ctypes binding of the C ABI in include/vt_api.h.

The library is the compiled HIP path; there is no fallback.  Importing
this module fails loudly when lib/libvoxeltoy.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libvoxeltoy.so")

VT_OK = 0

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Camera(C.Structure):
    _fields_ = [("inv_modelview", C.c_float * 16), ("proj", C.c_float * 16), ("inv_proj", C.c_float * 16),
                ("viewport", C.c_float * 4), ("win_z", C.c_float), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float), ("lens_model", C.c_int32)]


class Environment(C.Structure):
    _fields_ = [("use_image", C.c_int32), ("top", C.c_float * 3), ("bottom", C.c_float * 3),
                ("image_w", C.c_int32), ("image_h", C.c_int32), ("image_rgb", C.c_void_p),
                ("cdf_u_w", C.c_int32), ("cdf_u_h", C.c_int32), ("cdf_u", C.c_void_p),
                ("cdf_v_n", C.c_int32), ("cdf_v", C.c_void_p), ("integral", C.c_float),
                ("rotation_radians", C.c_float)]


class Settings(C.Structure):
    _fields_ = [("max_bounces", C.c_int32), ("exposure", C.c_float), ("gamma", C.c_float),
                ("selected_voxel", C.c_int32 * 3), ("wireframe_opacity", C.c_float),
                ("wireframe_thickness", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("visits", C.c_uint64), ("interactions", C.c_uint64),
                ("samples", C.c_uint64), ("env_samples", C.c_uint64), ("env_evals", C.c_uint64),
                ("kernel_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("render_mode", C.c_int32), ("dda_budget", C.c_int32), ("refill_min", C.c_int32),
                ("refill_chunk", C.c_int32), ("max_paths", C.c_uint64), ("ms", C.c_double),
                ("candidates", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CacheStats(C.Structure):
    _fields_ = [("queries", C.c_uint64), ("samples", C.c_uint64), ("steps", C.c_uint32), ("last_loss", C.c_float),
                ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchTuning(C.Structure):
    _fields_ = [("solo_paths", C.c_uint64), ("max_paths", C.c_uint64), ("dda_budget", C.c_int32),
                ("refill_min", C.c_int32), ("refill_chunk", C.c_int32), ("ms", C.c_double), ("candidates", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# symbol -> (restype, argtypes)
SIGNATURES = {
    "vt_last_error": (C.c_char_p, []),
    "vt_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "vt_destroy": (None, [C.c_void_p]),
    "vt_synchronize": (C.c_int, [C.c_void_p]),
    "vt_device_cus": (C.c_int, [C.c_void_p]),
    "vt_set_volume": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, i32p, C.c_void_p, C.c_int32,
                                C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "vt_set_volume_u8": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, u8p, i32p, C.c_void_p,
                                   C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "vt_get_bounds": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "vt_read_volume": (C.c_int, [C.c_void_p, i32p]),
    "vt_update_material": (C.c_int, [C.c_void_p, C.c_int32, f32p, C.c_int32]),
    "vt_read_materials": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "vt_set_camera": (C.c_int, [C.c_void_p, C.POINTER(Camera)]),
    "vt_set_environment": (C.c_int, [C.c_void_p, C.POINTER(Environment)]),
    "vt_set_settings": (C.c_int, [C.c_void_p, C.POINTER(Settings)]),
    "vt_set_frame": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "vt_reset_accum": (C.c_int, [C.c_void_p]),
    "vt_set_stats": (C.c_int, [C.c_void_p, C.c_int32]),
    "vt_set_render_mode": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64]),
    "vt_set_traversal_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "vt_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    "vt_batch_render": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                  C.c_uint32, C.c_int32]),
    "vt_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "vt_batch_destroy": (None, [C.c_void_p]),
    "vt_batch_set_tuning": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int32, C.c_int32, C.c_int32]),
    "vt_batch_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(BatchTuning)]),
    "vt_autotune_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.c_uint32, C.POINTER(BatchTuning)]),
    "vt_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Tuning)]),
    "vt_cache_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "vt_cache_destroy": (None, [C.c_void_p]),
    "vt_cache_num_params": (C.c_int32, []),
    "vt_cache_get_params": (C.c_int, [C.c_void_p, f32p]),
    "vt_cache_set_params": (C.c_int, [C.c_void_p, f32p]),
    "vt_cache_infer": (C.c_int, [C.c_void_p, C.c_int64, f32p, f32p]),
    "vt_cache_gradient": (C.c_int, [C.c_void_p, C.c_int64, f32p, f32p, f32p, C.POINTER(C.c_float)]),
    "vt_cache_train": (C.c_int, [C.c_void_p, C.c_int64, f32p, f32p, C.c_float, C.POINTER(C.c_float)]),
    "vt_render_cached": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32,
                                   C.c_int32, C.c_float, C.c_uint32, C.POINTER(CacheStats)]),
    "vt_render": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    "vt_read_accum": (C.c_int, [C.c_void_p, f32p]),
    "vt_read_sum64": (C.c_int, [C.c_void_p, f64p]),
    "vt_write_accum": (C.c_int, [C.c_void_p, f32p]),
    "vt_write_sum64": (C.c_int, [C.c_void_p, f64p]),
    "vt_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "vt_set_launch_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "vt_set_overlap": (C.c_int, [C.c_void_p, C.c_int32]),
    "vt_get_launch_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_double)]),
    "vt_trace_first_hit": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, i32p]),
    "vt_trace_rays": (C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, i32p]),
    "vt_trace_rays_noalloc": (C.c_int, [C.c_void_p, C.c_int32, f32p, f32p, i32p]),
    "vt_voxelize": (C.c_int, [C.c_void_p, f32p, C.c_int64, u32p, C.c_int64, f32p, C.c_int32, C.c_int32,
                              C.c_void_p]),
    "vt_camera_matrices": (C.c_int, [f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                     C.c_int32, f32p, f32p, f32p, f32p]),
    "vt_env_build_cdf": (C.c_int, [f32p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.c_void_p, C.POINTER(C.c_int32), C.c_void_p, C.POINTER(C.c_float)]),
    "vt_prune_emissive": (C.c_int32, [i32p, C.c_int32, C.c_int32, C.c_int32, i32p, C.c_int32]),
    "vt_mesh_transform": (C.c_int, [f32p, f32p, i32p, f32p]),
    "vt_device_math": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f32p, f32p, f32p]),
    "vt_copy_probe": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(C.c_double)]),
}


class VTError(RuntimeError):
    pass


def library_path() -> str:
    return _LIB_PATH


def load_library() -> C.CDLL:
    if not os.path.exists(_LIB_PATH):
        raise VTError(f"{_LIB_PATH} is missing: build it with `python voxeltoy_amd/build.py` "
                      "(or __graft_entry__.build()); there is no non-HIP fallback")
    lib = C.CDLL(_LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _check(code: int):
    if code != VT_OK:
        raise VTError(f"vt error {code}: {lib().vt_last_error().decode()}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def camera_matrices(eye, target, fov_y, near, far, lens_model, width, height):
    """Renderer::updateCamera matrices: returns mv, inv_mv, proj, inv_proj (4x4 row-major)."""
    out = [np.zeros(16, np.float32) for _ in range(4)]
    _check(lib().vt_camera_matrices(np.asarray(eye, np.float32), np.asarray(target, np.float32), float(fov_y),
                                    float(near), float(far), int(lens_model), int(width), int(height), *out))
    return [o.reshape(4, 4) for o in out]


def env_build_cdf(rgb: np.ndarray):
    """calculateCDF over an (h, w, 3) float image: returns cdf_u (h', w'+1), cdf_v (h'+1,), integral."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    cuw, cuh, cvn = C.c_int32(), C.c_int32(), C.c_int32()
    _check(lib().vt_env_build_cdf(rgb.reshape(-1), w, h, C.byref(cuw), C.byref(cuh), None, C.byref(cvn), None, None))
    cu = np.zeros(cuw.value * cuh.value, np.float32)
    cv = np.zeros(cvn.value, np.float32)
    integ = C.c_float()
    _check(lib().vt_env_build_cdf(rgb.reshape(-1), w, h, C.byref(cuw), C.byref(cuh), _ptr(cu), C.byref(cvn),
                                  _ptr(cv), C.byref(integ)))
    return cu.reshape(cuh.value, cuw.value), cv, integ.value


def mesh_transform(bmin, bmax, res) -> np.ndarray:
    """computeMeshTransform (src/renderer/import.cpp:46-64), row-major 4x4."""
    M = np.zeros(16, np.float32)
    _check(lib().vt_mesh_transform(np.asarray(bmin, np.float32), np.asarray(bmax, np.float32),
                                   np.asarray(res, np.int32), M))
    return M.reshape(4, 4)


def prune_emissive(offsets: np.ndarray, emissive: np.ndarray) -> np.ndarray:
    """pruneInteriorEmissiveVoxels; offsets is (nz, ny, nx) int32."""
    nz, ny, nx = offsets.shape
    lst = np.ascontiguousarray(emissive, np.int32).copy()
    n = lib().vt_prune_emissive(np.ascontiguousarray(offsets, np.int32).reshape(-1), nx, ny, nz, lst, len(lst))
    if n < 0:
        _check(n)
    return lst[:n]


class Context:
    """One device context of the tracer (owns all device memory)."""

    def __init__(self, device: int = 0):
        self._c = C.c_void_p()
        _check(lib().vt_create(device, C.byref(self._c)))
        self.width = 0
        self.height = 0
        self.shape = None
        self._fp64 = False

    def close(self):
        if self._c:
            lib().vt_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def cus(self) -> int:
        return lib().vt_device_cus(self._c)

    def synchronize(self):
        _check(lib().vt_synchronize(self._c))

    # -- volume -------------------------------------------------------------
    def set_volume(self, offsets: np.ndarray, mat: np.ndarray, emissive=None, bmin=None, bmax=None):
        """offsets: (nz, ny, nx) int32 material offsets (-1 empty)."""
        offsets = np.ascontiguousarray(offsets, np.int32)
        nz, ny, nx = offsets.shape
        mat = np.ascontiguousarray(mat, np.float32)
        em = None if emissive is None or len(emissive) == 0 else np.ascontiguousarray(emissive, np.int32)
        bmn = None if bmin is None else np.ascontiguousarray(bmin, np.float32)
        bmx = None if bmax is None else np.ascontiguousarray(bmax, np.float32)
        _check(lib().vt_set_volume(self._c, nx, ny, nz, offsets.reshape(-1), _ptr(mat), len(mat), _ptr(em),
                                   0 if em is None else len(em), _ptr(bmn), _ptr(bmx)))
        self.shape = (nz, ny, nx)

    def set_volume_u8(self, ids: np.ndarray, id_offset: np.ndarray, mat: np.ndarray, emissive=None, bmin=None,
                      bmax=None):
        ids = np.ascontiguousarray(ids, np.uint8)
        nz, ny, nx = ids.shape
        mat = np.ascontiguousarray(mat, np.float32)
        em = None if emissive is None or len(emissive) == 0 else np.ascontiguousarray(emissive, np.int32)
        bmn = None if bmin is None else np.ascontiguousarray(bmin, np.float32)
        bmx = None if bmax is None else np.ascontiguousarray(bmax, np.float32)
        _check(lib().vt_set_volume_u8(self._c, nx, ny, nz, ids.reshape(-1),
                                      np.ascontiguousarray(id_offset, np.int32), _ptr(mat), len(mat), _ptr(em),
                                      0 if em is None else len(em), _ptr(bmn), _ptr(bmx)))
        self.shape = (nz, ny, nx)

    def bounds(self):
        bmin, bmax, vs = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
        _check(lib().vt_get_bounds(self._c, bmin, bmax, vs))
        return bmin, bmax, vs

    def read_volume(self) -> np.ndarray:
        out = np.zeros(self.shape, np.int32)
        _check(lib().vt_read_volume(self._c, out.reshape(-1)))
        return out

    def update_material(self, offset: int, values):
        v = np.ascontiguousarray(values, np.float32)
        _check(lib().vt_update_material(self._c, int(offset), v, len(v)))

    def read_materials(self) -> np.ndarray:
        n = C.c_int32()
        _check(lib().vt_read_materials(self._c, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(lib().vt_read_materials(self._c, _ptr(out), n.value, C.byref(n)))
        return out

    # -- state --------------------------------------------------------------
    def set_camera(self, inv_mv, proj, inv_proj, viewport, win_z, lens_model=0, lens_radius=0.0,
                   focal_distance=99999999.0):
        cam = Camera()
        cam.inv_modelview[:] = np.asarray(inv_mv, np.float32).reshape(-1)
        cam.proj[:] = np.asarray(proj, np.float32).reshape(-1)
        cam.inv_proj[:] = np.asarray(inv_proj, np.float32).reshape(-1)
        cam.viewport[:] = np.asarray(viewport, np.float32)
        cam.win_z = win_z
        cam.lens_model = lens_model
        cam.lens_radius = lens_radius
        cam.focal_distance = focal_distance
        _check(lib().vt_set_camera(self._c, C.byref(cam)))

    def set_environment(self, top=(1, 1, 1), bottom=(1, 1, 1), image=None, cdf_u=None, cdf_v=None, integral=0.0,
                        rotation=0.0):
        env = Environment()
        env.top[:] = list(np.asarray(top, np.float32))
        env.bottom[:] = list(np.asarray(bottom, np.float32))
        env.rotation_radians = rotation
        keep = []
        if image is not None:
            img = np.ascontiguousarray(image, np.float32)
            cu = np.ascontiguousarray(cdf_u, np.float32)
            cv = np.ascontiguousarray(cdf_v, np.float32)
            # the library copies h*w*3, rows*cols and n floats from these
            if img.ndim != 3 or img.shape[2] != 3:
                raise ValueError(f"environment image must be (h, w, 3), got {img.shape}")
            if cu.ndim != 2:
                raise ValueError(f"cdf_u must be 2-D (h', w'+1), got {cu.shape}")
            if cv.ndim != 1:
                raise ValueError(f"cdf_v must be 1-D (h'+1,), got {cv.shape}")
            keep = [img, cu, cv]
            env.use_image = 1
            env.image_h, env.image_w = img.shape[:2]
            env.image_rgb = img.ctypes.data
            env.cdf_u_h, env.cdf_u_w = cu.shape
            env.cdf_u = cu.ctypes.data
            env.cdf_v_n = len(cv)
            env.cdf_v = cv.ctypes.data
            env.integral = integral
        _check(lib().vt_set_environment(self._c, C.byref(env)))
        del keep

    def set_settings(self, max_bounces=1, exposure=1.0, gamma=2.2, selected=(-(2 ** 31),) * 3,
                     wireframe_opacity=0.0, wireframe_thickness=0.01):
        st = Settings()
        st.max_bounces = max_bounces
        st.exposure = exposure
        st.gamma = gamma
        st.selected_voxel[:] = list(selected)
        st.wireframe_opacity = wireframe_opacity
        st.wireframe_thickness = wireframe_thickness
        _check(lib().vt_set_settings(self._c, C.byref(st)))

    def set_frame(self, width: int, height: int, fp64: bool = False):
        _check(lib().vt_set_frame(self._c, width, height, 1 if fp64 else 0))
        self.width, self.height, self._fp64 = width, height, fp64

    def reset_accum(self):
        _check(lib().vt_reset_accum(self._c))

    def set_stats(self, enabled: bool):
        _check(lib().vt_set_stats(self._c, 1 if enabled else 0))

    def set_render_mode(self, mode: int, max_paths: int = 0):
        """0 = megakernel, 1 = wavefront (default); both give identical images."""
        _check(lib().vt_set_render_mode(self._c, mode, max_paths))

    def set_traversal_tuning(self, dda_budget: int = 0, refill_min: int = 0, refill_chunk: int = 0):
        """Wavefront traversal: DDA steps between refill checks, idle lanes
        per wave that trigger a refill, queue entries reserved per atomic
        (-1 = exactly the idle lanes); 0 = default.  Images are identical."""
        _check(lib().vt_set_traversal_tuning(self._c, dda_budget, refill_min, refill_chunk))

    def autotune(self, passes: int = 1) -> dict:
        """Time every kernel organisation / traversal tuning on `passes`
        passes of the current scene, keep the fastest (images are identical)."""
        t = Tuning()
        _check(lib().vt_autotune(self._c, passes, C.byref(t)))
        return t.as_dict()

    # -- rendering ----------------------------------------------------------
    def render(self, seed: int, n0: int, count: int, rng_frame_max: int = 0xFFFFFFFF):
        _check(lib().vt_render(self._c, seed, n0, count, rng_frame_max))

    def read_accum(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        _check(lib().vt_read_accum(self._c, out.reshape(-1)))
        return out

    def read_sum64(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float64)
        _check(lib().vt_read_sum64(self._c, out.reshape(-1)))
        return out

    def write_accum(self, rgba: np.ndarray):
        """Resume: the running mean read earlier with read_accum."""
        a = np.ascontiguousarray(rgba, np.float32)
        if a.size != self.height * self.width * 4:
            raise ValueError("running mean of another frame size")
        _check(lib().vt_write_accum(self._c, a.reshape(-1)))

    def write_sum64(self, rgb: np.ndarray):
        a = np.ascontiguousarray(rgb, np.float64)
        if a.size != self.height * self.width * 3:
            raise ValueError("FP64 sums of another frame size")
        _check(lib().vt_write_sum64(self._c, a.reshape(-1)))

    def stats(self) -> dict:
        s = Stats()
        _check(lib().vt_get_stats(self._c, C.byref(s)))
        return s.as_dict()

    def set_overlap(self, streams: int):
        """1 or 2 streams for wavefront renders of many passes (identical images)."""
        _check(lib().vt_set_overlap(self._c, streams))

    def set_launch_timing(self, enabled: bool):
        """Bracket every wavefront traversal launch with HIP events."""
        _check(lib().vt_set_launch_timing(self._c, 1 if enabled else 0))

    def launch_timing(self) -> dict:
        """Summed traversal-kernel time (ms), launch count and busy time
        (union of the launches' intervals) of the last render."""
        ms, n, u = C.c_double(), C.c_uint32(), C.c_double()
        _check(lib().vt_get_launch_timing(self._c, C.byref(ms), C.byref(n), C.byref(u)))
        return {"traverse_ms": ms.value, "launches": n.value, "union_ms": u.value}

    def first_hit(self, seed: int, frame: int) -> np.ndarray:
        out = np.zeros((self.height, self.width, 8), np.int32)
        _check(lib().vt_trace_first_hit(self._c, seed, frame, out.reshape(-1)))
        return out

    def trace_rays(self, origins: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        o = np.ascontiguousarray(origins, np.float32).reshape(-1)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1)
        n = len(o) // 3
        out = np.zeros((n, 8), np.int32)
        _check(lib().vt_trace_rays(self._c, n, o, d, out.reshape(-1)))
        return out

    def trace_rays_noalloc(self, origins: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        """trace_rays through the library's fixed per-device staging array
        (no device allocation; at most VT_NOALLOC_MAX_RAYS rays)."""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1)
        n = len(o) // 3
        out = np.zeros((n, 8), np.int32)
        _check(lib().vt_trace_rays_noalloc(self._c, n, o, d, out.reshape(-1)))
        return out

    def voxelize(self, verts: np.ndarray, tris: np.ndarray, M: np.ndarray, mode: int = 0, material_offset: int = 0,
                 want_occupancy: bool = True):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1)
        t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
        occ = np.zeros(self.shape, np.uint8) if want_occupancy else None
        _check(lib().vt_voxelize(self._c, v, len(v) // 3, t, len(t) // 3,
                                 np.ascontiguousarray(M, np.float32).reshape(-1), mode, material_offset, _ptr(occ)))
        return occ

    def copy_probe(self, nbytes: int, iters: int = 10) -> float:
        """Mean kernel ms of a streaming device copy of nbytes."""
        ms = C.c_double()
        _check(lib().vt_copy_probe(self._c, nbytes, iters, C.byref(ms)))
        return ms.value

    def device_math(self, fn: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        out = np.zeros_like(a)
        _check(lib().vt_device_math(self._c, fn, len(a), a, b, out))
        return out


class Batch:
    """Several contexts (scenes) rendered by one wavefront launch, optionally
    through a recorded HIP graph.  The contexts must outlive the batch."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c._c.value for c in self.contexts])
        self._b = C.c_void_p()
        _check(lib().vt_batch_create(arr, len(self.contexts), C.byref(self._b)))

    def render(self, seeds, counts, n0=None, rng_frame_max: int = 0xFFFFFFFF, graph: bool = False,
               stats: bool = False):
        n = len(self.contexts)
        s = (C.c_uint64 * n)(*[int(x) for x in seeds])
        k = (C.c_uint32 * n)(*[int(x) for x in counts])
        z = None if n0 is None else (C.c_uint32 * n)(*[int(x) for x in n0])
        _check(lib().vt_batch_render(self._b, s, z, k, rng_frame_max, (1 if graph else 0) | (2 if stats else 0)))

    def stats(self) -> dict:
        st = Stats()
        _check(lib().vt_batch_get_stats(self._b, C.byref(st)))
        return st.as_dict()

    def set_tuning(self, solo_paths: int = 0, max_paths: int = 0, dda_budget: int = 0, refill_min: int = 0,
                   refill_chunk: int = 0):
        _check(lib().vt_batch_set_tuning(self._b, solo_paths, max_paths, dda_budget, refill_min, refill_chunk))

    def tuning(self) -> dict:
        t = BatchTuning()
        _check(lib().vt_batch_get_tuning(self._b, C.byref(t)))
        return t.as_dict()

    def autotune(self, seeds, counts, n0=None, rng_frame_max: int = 0xFFFFFFFF) -> dict:
        """Time graph replays of this render over the batch's tunings and
        keep the fastest (the scenes' accumulations are restored)."""
        n = len(self.contexts)
        s = (C.c_uint64 * n)(*[int(x) for x in seeds])
        k = (C.c_uint32 * n)(*[int(x) for x in counts])
        z = None if n0 is None else (C.c_uint32 * n)(*[int(x) for x in n0])
        t = BatchTuning()
        _check(lib().vt_autotune_batch(self._b, s, z, k, rng_frame_max, C.byref(t)))
        return t.as_dict()

    def close(self):
        if self._b:
            lib().vt_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RadianceCache:
    """The learned radiance cache of one context (see vt_api.h)."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        self._p = C.c_void_p()
        _check(lib().vt_cache_create(ctx._c, seed, C.byref(self._p)))
        self.n_params = lib().vt_cache_num_params()

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(lib().vt_cache_get_params(self._p, p))
        return p

    def set_params(self, p):
        _check(lib().vt_cache_set_params(self._p, np.ascontiguousarray(p, np.float32)))

    def infer(self, raw) -> np.ndarray:
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 12)
        out = np.zeros((len(raw), 3), np.float32)
        _check(lib().vt_cache_infer(self._p, len(raw), raw.reshape(-1), out.reshape(-1)))
        return out

    def gradient(self, raw, target):
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 12)
        tgt = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        _check(lib().vt_cache_gradient(self._p, len(raw), raw.reshape(-1), tgt.reshape(-1), g, C.byref(loss)))
        return g, loss.value

    def train(self, raw, target, lr: float = 1e-3) -> float:
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 12)
        tgt = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = C.c_float()
        _check(lib().vt_cache_train(self._p, len(raw), raw.reshape(-1), tgt.reshape(-1), lr, C.byref(loss)))
        return loss.value

    def render(self, seed: int, n0: int, count: int, cache_bounce: int = 2, train_paths: int = 16384,
               lr: float = 3e-3, batch: int = 16384, train_extra: int = -1) -> dict:
        """`count` passes of the context's frame with paths ending into the
        cache at bounce `cache_bounce` (training paths `train_extra` bounces
        later, -1 = never), training after every pass."""
        st = CacheStats()
        _check(lib().vt_render_cached(self.ctx._c, self._p, seed, n0, count, cache_bounce, train_paths, train_extra,
                                      lr, batch, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self._p:
            lib().vt_cache_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
