"""Host-side mirror of sprout's triangle BVH interface, running on the GPU.

sprout (core/scene/shape/triangle/bvh):
    Builder_SAH(num_slices, sweep_threshold, max_primitives)
        .build(tree, num_triangles, triangles, vertices, threads)   triangle_bvh_builder_sah.cpp:21
    Tree.intersect(ray_origin, ray_direction, min_t, max_t&, nodes, isec)  triangle_bvh_tree.cpp:41
    Tree.intersect_p(ray_origin, ray_direction, min_t, max_t, nodes)       triangle_bvh_tree.cpp:145

Here a Tree lives in HBM and is queried with whole batches of rays at once
(device tensors); the per-ray arithmetic, visiting order and tie rule are the
reference's.  Rays are (n, 8) float32 {origin.xyz, min_t, direction.xyz,
max_t}; closest hits are (n, 4) 32-bit words {t, u, v, input triangle or
0xFFFFFFFF}.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import MI_NULL, MiBvhInfo, MiBvhSettings, check

NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("data", "<u4"), ("hi", "<f4", 3), ("axis", "u1"),
                       ("num", "u1"), ("pad", "u1", 2)])

_context_device = None


def init(device: int = 0) -> None:
    """mi_init on `device` (idempotent for the same device)."""
    global _context_device
    if _context_device == device:
        return
    check(_lib.load().mi_init(device), "mi_init")
    _context_device = device


def release() -> None:
    global _context_device
    if _context_device is not None:
        check(_lib.load().mi_release(), "mi_release")
        _context_device = None


def _stream(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def _dev(t: torch.Tensor, dtype, cols=None) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"expected shape (n, {cols}), got {tuple(t.shape)}")
    return t


class Tree:
    """A triangle SBVH resident on the device (a mi_bvh handle)."""

    def __init__(self, handle: int):
        self.handle = handle

    def __del__(self):
        if getattr(self, "handle", None) is not None and _context_device is not None:
            _lib.load().mi_bvh_release(self.handle)
            self.handle = None

    def info(self) -> MiBvhInfo:
        info = MiBvhInfo()
        check(_lib.load().mi_bvh_get_info(self.handle, ctypes.byref(info)), "mi_bvh_get_info")
        return info

    @property
    def num_nodes(self) -> int:
        return self.info().num_nodes

    def download(self):
        """(nodes structured array in sprout's 32-byte layout, slots uint32)."""
        info = self.info()
        nodes = np.zeros(info.num_nodes, NODE_DTYPE)
        slots = np.zeros(info.num_slots, np.uint32)
        check(_lib.load().mi_bvh_download(self.handle, nodes.ctypes.data, slots.ctypes.data),
              "mi_bvh_download")
        return nodes, slots

    def intersect(self, rays: torch.Tensor, hits: torch.Tensor | None = None, stream=None):
        """Closest hits (Tree::intersect) for a (n, 8) float32 device tensor of rays."""
        rays = _dev(rays, torch.float32, 8)
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        check(_lib.load().mi_trace_closest(self.handle, rays.data_ptr(), n, hits.data_ptr(),
                                           _stream(stream)), "mi_trace_closest")
        return hits

    def intersect_p(self, rays: torch.Tensor, occluded: torch.Tensor | None = None, stream=None):
        """Occlusion flags (Tree::intersect_p) as an (n,) int32 device tensor."""
        rays = _dev(rays, torch.float32, 8)
        n = rays.shape[0]
        if occluded is None:
            occluded = torch.empty((n,), dtype=torch.int32, device=rays.device)
        check(_lib.load().mi_trace_any(self.handle, rays.data_ptr(), n, occluded.data_ptr(),
                                       _stream(stream)), "mi_trace_any")
        return occluded

    def counts(self, rays: torch.Tensor, any_hit: bool = False, stream=None):
        """(results, counts): counts[i] = (nodes fetched, triangle tests) of the reference loop."""
        rays = _dev(rays, torch.float32, 8)
        n = rays.shape[0]
        out = torch.empty((n,) if any_hit else (n, 4), dtype=torch.int32, device=rays.device)
        cnt = torch.empty((n, 2), dtype=torch.int32, device=rays.device)
        check(_lib.load().mi_trace_counts(self.handle, rays.data_ptr(), n, int(any_hit),
                                          out.data_ptr(), cnt.data_ptr(), _stream(stream)),
              "mi_trace_counts")
        return out, cnt


class Builder_SAH:
    """SBVH builder with sprout's settings (triangle_mesh_provider.cpp:377: 16, 64, 4)."""

    def __init__(self, num_slices: int = 16, sweep_threshold: int = 64, max_primitives: int = 4):
        self.settings = MiBvhSettings(num_slices, sweep_threshold, max_primitives)

    def build(self, positions: torch.Tensor, indices: torch.Tensor, stream=None) -> Tree:
        """Builds on the device from (v, 3) float32 positions and (t, 3) int32 indices."""
        positions = _dev(positions, torch.float32, 3)
        indices = _dev(indices, torch.int32, 3)
        handle = ctypes.c_uint32(MI_NULL)
        check(_lib.load().mi_bvh_build(positions.data_ptr(), positions.shape[0], 3,
                                       indices.data_ptr(), indices.shape[0],
                                       ctypes.byref(self.settings), _stream(stream),
                                       ctypes.byref(handle)), "mi_bvh_build")
        return Tree(handle.value)


def upload(nodes: np.ndarray, slots: np.ndarray, positions: np.ndarray, indices: np.ndarray) -> Tree:
    """Uploads a host-built tree in sprout's serialized layout (mi_bvh_upload)."""
    nodes = np.ascontiguousarray(nodes)
    assert nodes.dtype.itemsize == 32
    slots = np.ascontiguousarray(slots, np.uint32)
    positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    indices = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    handle = ctypes.c_uint32(MI_NULL)
    check(_lib.load().mi_bvh_upload(nodes.ctypes.data, len(nodes), slots.ctypes.data, len(slots),
                                    positions.ctypes.data, len(positions), 3, indices.ctypes.data,
                                    len(indices), ctypes.byref(handle)), "mi_bvh_upload")
    return Tree(handle.value)


def set_trace_config(block_size: int = 0, blocks_per_cu: int = 0, kernel: int = 0,
                     refill_lanes: int = 0, claim_batch: int = 0, lds_entries: int = 0,
                     sort_rays: int = 0) -> None:
    """Traversal launch shape; 0 keeps the library default of each field
    (sort_rays: 0 off)."""
    cfg = _lib.MiTraceConfig(block_size, blocks_per_cu, kernel, refill_lanes, claim_batch,
                             lds_entries, sort_rays)
    check(_lib.load().mi_set_trace_config(ctypes.byref(cfg)), "mi_set_trace_config")


def autotune(tree, rays, repeats: int = 5):
    """mi_trace_autotune over device rays (n, 8) float32: makes the fastest
    launch shape the context's; returns (shape as a dict, mean ms)."""
    import torch

    r = rays if isinstance(rays, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    cfg = _lib.MiTraceConfig()
    ms = ctypes.c_float(0)
    check(_lib.load().mi_trace_autotune(tree.handle, ctypes.c_void_p(r.data_ptr()), r.shape[0], repeats,
                                        _stream(None), ctypes.byref(cfg), ctypes.byref(ms)),
          "mi_trace_autotune")
    return {f: getattr(cfg, f) for f, _ in cfg._fields_}, ms.value


def sync(stream=None) -> None:
    check(_lib.load().mi_sync(_stream(stream)), "mi_sync")
