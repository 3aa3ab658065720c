"""Build the voxel path tracer library (libvoxeltoy.so) for gfx950.

HIP sources are compiled with hipcc, host C++ with g++; both without FMA
contraction so that the float arithmetic of the device kernels and the host
preparation steps is exactly the arithmetic written in the source.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libvoxeltoy.so")

HIP_SOURCES = ["csrc/vt_trace.hip", "csrc/vt_volume.hip", "csrc/vt_cache.hip", "csrc/vt_api.hip"]
CXX_SOURCES = ["host/vt_host_algos.cpp", "host/camera.cpp", "host/scene_io.cpp", "host/renderer.cpp",
               "host/cpu_voxelizer.cpp"]
HEADERS = ["csrc/vtmath.h", "csrc/vt_device.h", "csrc/vt_kernels.h", "csrc/vt_cache.h", "host/vt_host_algos.h",
           "host/camera.h", "host/scene_io.h", "host/renderer.h", "host/vt_math_types.h"]
API_HEADERS = ["vt_api.h", "vt_renderer.h"]

ARCH = os.environ.get("VT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "host")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def build(force: bool = True, jobs: int = 8, verbose: bool = False) -> str:
    """Compile every source and link lib/libvoxeltoy.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    existing_cxx = [s for s in CXX_SOURCES if os.path.exists(os.path.join(PKG, s))]
    hdr_mtime = max(os.path.getmtime(os.path.join(PKG, h)) for h in HEADERS if os.path.exists(os.path.join(PKG, h)))
    hdr_mtime = max([hdr_mtime] + [os.path.getmtime(os.path.join(ROOT, "include", h)) for h in API_HEADERS])
    cmds = []
    for s in HIP_SOURCES:
        src, obj = os.path.join(PKG, s), _obj(s)
        if force or not os.path.exists(obj) or os.path.getmtime(obj) < max(os.path.getmtime(src), hdr_mtime):
            cmds.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    for s in existing_cxx:
        src, obj = os.path.join(PKG, s), _obj(s)
        if force or not os.path.exists(obj) or os.path.getmtime(obj) < max(os.path.getmtime(src), hdr_mtime):
            cmds.append([CXX, *CXX_FLAGS, "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    objs = [_obj(s) for s in HIP_SOURCES] + [_obj(s) for s in existing_cxx]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    path = build(force="--incremental" not in sys.argv, verbose="-v" in sys.argv)
    print(path)
