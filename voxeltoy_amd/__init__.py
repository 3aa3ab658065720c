"""MI355X-native voxel path tracer: voxelToy's rendering hot path as HIP kernels
for gfx950 behind a C ABI (include/vt_api.h)."""
from . import api, scenes, workloads  # noqa: F401
from .api import Context, VTError  # noqa: F401

__all__ = ["api", "scenes", "workloads", "Context", "VTError"]
