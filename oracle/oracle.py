"""ctypes binding of the CPU oracle (oracle/build/libvtoracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py as the checker; never by the product.

The oracle is a C restatement of voxelToy's GLSL hot path (see
vt_oracle.c).  Parity status of each function is recorded in its
docstring here and pinned by tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libvtoracle.so")


def build(force: bool = False) -> str:
    """make the library; one build at a time (test workers call this
    concurrently), under a lock file next to it."""
    import fcntl
    os.makedirs(os.path.join(HERE, "build"), exist_ok=True)
    args = ["make", "-C", HERE]
    if force:
        args.insert(1, "-B")
    with open(os.path.join(HERE, "build", ".lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        subprocess.run(args, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return LIB_PATH


class Scene(C.Structure):
    _fields_ = [("res", C.c_int32 * 3), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("vsize", C.c_float * 3),
                ("offsets", C.c_void_p), ("mat", C.c_void_p), ("mat_size", C.c_int32), ("emissive", C.c_void_p),
                ("n_emissive", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("inv_mv", C.c_float * 16), ("proj", C.c_float * 16), ("inv_proj", C.c_float * 16),
                ("viewport", C.c_float * 4), ("win_z", C.c_float), ("lens_radius", C.c_float),
                ("focal_distance", C.c_float), ("lens_model", C.c_int32)]


class Env(C.Structure):
    _fields_ = [("use_image", C.c_int32), ("top", C.c_float * 3), ("bottom", C.c_float * 3), ("img_w", C.c_int32),
                ("img_h", C.c_int32), ("img", C.c_void_p), ("cdfu_w", C.c_int32), ("cdfu_h", C.c_int32),
                ("cdfu", C.c_void_p), ("cdfv_n", C.c_int32), ("cdfv", C.c_void_p), ("integral", C.c_float),
                ("rotation", C.c_float)]


class Settings(C.Structure):
    _fields_ = [("max_bounces", C.c_int32), ("exposure", C.c_float), ("gamma", C.c_float),
                ("selected", C.c_int32 * 3), ("wire_opacity", C.c_float), ("wire_thickness", C.c_float)]


class Counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("visits", C.c_uint64), ("interactions", C.c_uint64),
                ("samples", C.c_uint64), ("env_samples", C.c_uint64), ("env_evals", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
P = C.POINTER

_SIG = {
    "orc_aabb": (C.c_float, [f32p, f32p, f32p, f32p]),
    "orc_max_steps": (C.c_int, [i32p]),
    "orc_raymarch": (C.c_int, [P(Scene), f32p, f32p, C.c_int, i32p, P(C.c_int32), P(C.c_int32)]),
    "orc_traverse": (C.c_int, [P(Scene), f32p, f32p, i32p, P(C.c_int32), P(C.c_int32)]),
    "orc_basis": (None, [P(Scene), i32p, f32p, f32p, f32p]),
    "orc_generate_ray": (None, [P(Camera), C.c_float, C.c_float, f32p, f32p, f32p]),
    "orc_pcg32_raw": (None, [C.c_uint64, C.c_uint64, C.c_int, u32p]),
    "orc_pcg32_stream": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, f32p]),
    "orc_env_sample_uv": (None, [P(Env), C.c_float, C.c_float, f32p, i32p]),
    "orc_env_texture": (None, [P(Env), C.c_float, C.c_float, f32p]),
    "orc_uv_from_dir": (None, [f32p, C.c_float, f32p]),
    "orc_powf": (C.c_float, [C.c_float, C.c_float]),
    "orc_sincosf": (None, [C.c_float, P(C.c_float), P(C.c_float)]),
    "orc_atan2f": (C.c_float, [C.c_float, C.c_float]),
    "orc_acosf": (C.c_float, [C.c_float]),
    "orc_log2f": (C.c_float, [C.c_float]),
    "orc_exp2f": (C.c_float, [C.c_float]),
    "orc_first_hit_image": (None, [P(Scene), P(Camera), C.c_int32, C.c_int32, C.c_uint64, C.c_uint32, i32p]),
    "orc_trace_rays": (None, [P(Scene), C.c_int32, f32p, f32p, i32p]),
    "orc_sample": (None, [P(Scene), P(Env), P(Camera), P(Settings), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                          C.c_uint64, C.c_uint32, C.c_uint32, f32p, P(Counters)]),
    "orc_render": (None, [P(Scene), P(Env), P(Camera), P(Settings), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                          C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, f32p, C.c_void_p, P(Counters),
                          C.c_int32]),
    "orc_voxelize": (None, [f32p, u32p, C.c_int64, f32p, i32p, C.c_int32, C.c_uint8, u8p]),
    "orc_mesh_transform": (None, [f32p, f32p, i32p, f32p]),
    "orc_calculate_cdf": (None, [f32p, C.c_int32, C.c_int32, f32p, f32p, P(C.c_float)]),
    "orc_image_function": (None, [f32p, C.c_int32, C.c_int32, C.c_int32, f32p, P(C.c_int32), P(C.c_int32)]),
    "orc_prune_emissive": (C.c_int32, [i32p, i32p, i32p, C.c_int32]),
    "orc_camera_matrices": (None, [f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32, C.c_int32,
                                   f32p, f32p, f32p, f32p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def f32(x):
    return np.ascontiguousarray(x, np.float32)


class OracleScene:
    """Holds the arrays behind an orc_scene / orc_env / orc_camera."""

    def __init__(self, offsets: np.ndarray, mat: np.ndarray, emissive=None, bmin=None, bmax=None):
        self.offsets = np.ascontiguousarray(offsets, np.int32)
        nz, ny, nx = self.offsets.shape
        self.mat = f32(mat)
        self.emissive = np.ascontiguousarray(np.zeros(0, np.int32) if emissive is None else emissive, np.int32)
        if bmin is None:
            vs = np.float32(1000.0) / np.float32(max(nx, ny, nz))
            size = np.array([vs * np.float32(nx), vs * np.float32(ny), vs * np.float32(nz)], np.float32)
            bmin, bmax = -size * np.float32(0.5), size * np.float32(0.5)
        self.bmin, self.bmax = f32(bmin), f32(bmax)
        res = np.array([nx, ny, nz], np.float32)
        self.vsize = ((self.bmax - self.bmin) / res).astype(np.float32)
        s = Scene()
        s.res[:] = [nx, ny, nz]
        s.bmin[:] = list(self.bmin)
        s.bmax[:] = list(self.bmax)
        s.vsize[:] = list(self.vsize)
        s.offsets = self.offsets.ctypes.data
        s.mat = self.mat.ctypes.data
        s.mat_size = len(self.mat)
        s.emissive = self.emissive.ctypes.data if len(self.emissive) else None
        s.n_emissive = len(self.emissive)
        self.s = s
        self.env_keep = []
        self.env = Env()
        self.set_environment()
        self.st = Settings()
        self.set_settings()
        self.cam = Camera()

    def set_environment(self, top=(1, 1, 1), bottom=(1, 1, 1), image=None, cdf_u=None, cdf_v=None, integral=0.0,
                        rotation=0.0):
        e = Env()
        e.top[:] = list(f32(top))
        e.bottom[:] = list(f32(bottom))
        e.rotation = rotation
        self.env_keep = []
        if image is not None:
            img, cu, cv = f32(image), f32(cdf_u), f32(cdf_v)
            self.env_keep = [img, cu, cv]
            e.use_image = 1
            e.img_h, e.img_w = img.shape[:2]
            e.img = img.ctypes.data
            e.cdfu_h, e.cdfu_w = cu.shape
            e.cdfu = cu.ctypes.data
            e.cdfv_n = len(cv)
            e.cdfv = cv.ctypes.data
            e.integral = integral
        self.env = e

    def set_settings(self, max_bounces=1, exposure=1.0, gamma=2.2, selected=(-(2 ** 31),) * 3,
                     wireframe_opacity=0.0, wireframe_thickness=0.01):
        st = Settings()
        st.max_bounces = max_bounces
        st.exposure = exposure
        st.gamma = gamma
        st.selected[:] = list(selected)
        st.wire_opacity = wireframe_opacity
        st.wire_thickness = wireframe_thickness
        self.st = st

    def set_camera(self, inv_mv, proj, inv_proj, viewport, win_z, lens_model=0, lens_radius=0.0,
                   focal_distance=99999999.0):
        c = Camera()
        c.inv_mv[:] = list(f32(inv_mv).reshape(-1))
        c.proj[:] = list(f32(proj).reshape(-1))
        c.inv_proj[:] = list(f32(inv_proj).reshape(-1))
        c.viewport[:] = list(f32(viewport))
        c.win_z = win_z
        c.lens_model = lens_model
        c.lens_radius = lens_radius
        c.focal_distance = focal_distance
        self.cam = c

    # -- calls ----------------------------------------------------------------
    def first_hit(self, W, H, seed, frame):
        out = np.zeros((H, W, 8), np.int32)
        lib().orc_first_hit_image(C.byref(self.s), C.byref(self.cam), W, H, seed, frame, out.reshape(-1))
        return out

    def trace_rays(self, o, d):
        o, d = f32(o).reshape(-1), f32(d).reshape(-1)
        out = np.zeros((len(o) // 3, 8), np.int32)
        lib().orc_trace_rays(C.byref(self.s), len(o) // 3, o, d, out.reshape(-1))
        return out

    def render(self, W, H, seed, n0, count, rng_frame_max=0xFFFFFFFF, y0=0, y1=None, accum=None, sum64=None,
               threads=0):
        if y1 is None:
            y1 = H
        if accum is None:
            accum = np.zeros((H, W, 4), np.float32)
        cnt = Counters()
        lib().orc_render(C.byref(self.s), C.byref(self.env), C.byref(self.cam), C.byref(self.st), W, H, y0, y1,
                         seed, n0, count, rng_frame_max, accum.reshape(-1),
                         None if sum64 is None else sum64.ctypes.data_as(C.c_void_p), C.byref(cnt),
                         threads or os.cpu_count())
        return accum, cnt.as_dict()

    def sample(self, W, H, px, py, seed, frame, rng_frame=None):
        out = np.zeros(3, np.float32)
        cnt = Counters()
        lib().orc_sample(C.byref(self.s), C.byref(self.env), C.byref(self.cam), C.byref(self.st), W, H, px, py, seed,
                         frame, frame if rng_frame is None else rng_frame, out, C.byref(cnt))
        return out, cnt.as_dict()


def voxelize(verts, tris, M, res, mode=0, value=1):
    occ = np.zeros((res[2], res[1], res[0]), np.uint8)
    lib().orc_voxelize(f32(verts).reshape(-1), np.ascontiguousarray(tris, np.uint32).reshape(-1), len(tris),
                       f32(M).reshape(-1), np.asarray(res, np.int32), mode, value, occ.reshape(-1))
    return occ


def mesh_transform(bmin, bmax, res):
    M = np.zeros(16, np.float32)
    lib().orc_mesh_transform(f32(bmin), f32(bmax), np.asarray(res, np.int32), M)
    return M.reshape(4, 4)


def camera_matrices(eye, target, fov_y, near, far, lens_model, W, H):
    out = [np.zeros(16, np.float32) for _ in range(4)]
    lib().orc_camera_matrices(f32(eye), f32(target), fov_y, near, far, lens_model, W, H, *out)
    return [o.reshape(4, 4) for o in out]


def calculate_cdf(func: np.ndarray):
    h, w = func.shape
    cu = np.zeros((h, w + 1), np.float32)
    cv = np.zeros(h + 1, np.float32)
    integ = C.c_float()
    lib().orc_calculate_cdf(f32(func).reshape(-1), w, h, cu.reshape(-1), cv, C.byref(integ))
    return cu, cv, integ.value


def image_function(rgb: np.ndarray, max_size=512):
    h, w = rgb.shape[:2]
    out = np.zeros(w * h, np.float32)
    ow, oh = C.c_int32(), C.c_int32()
    lib().orc_image_function(f32(rgb).reshape(-1), w, h, max_size, out, C.byref(ow), C.byref(oh))
    return out[:ow.value * oh.value].reshape(oh.value, ow.value)


def build_env_cdf(rgb: np.ndarray):
    f = image_function(rgb)
    return calculate_cdf(f)


def prune_emissive(offsets: np.ndarray, emissive):
    nz, ny, nx = offsets.shape
    lst = np.ascontiguousarray(emissive, np.int32).copy()
    n = lib().orc_prune_emissive(np.ascontiguousarray(offsets, np.int32).reshape(-1), np.array([nx, ny, nz], np.int32),
                                 lst, len(lst))
    return lst[:n]


def pcg_raw(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_raw(initstate, initseq, n, out)
    return out


def pcg_stream(seed, pixel, frame, n):
    out = np.zeros(n, np.float32)
    lib().orc_pcg32_stream(seed, pixel, frame, n, out)
    return out


def math1(name, *args):
    return getattr(lib(), name)(*args)


def sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincosf(x, C.byref(s), C.byref(c))
    return s.value, c.value
