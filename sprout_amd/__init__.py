"""sprout_amd: MI355X-native SBVH build and ray traversal for the sprout renderer.

The compute lives in libsprout_gpu.so (HIP, gfx950) behind the C ABI of
include/sprout_gpu.h; this package is its ctypes front end.
"""

from ._lib import LIB_PATH, MI_NULL, MiError, load  # noqa: F401

__all__ = ["LIB_PATH", "MI_NULL", "MiError", "load"]
