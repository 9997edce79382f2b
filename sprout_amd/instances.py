"""Instanced meshes under a top-level tree: sprout's props
(core/scene/prop/prop_bvh_wrapper.cpp:33-247, core/scene/prop/prop.cpp:68-126,
core/scene/shape/triangle/triangle_mesh.cpp:366-446) on the GPU."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import MI_NULL, check
from .bvh import NODE_DTYPE, Tree, _dev, _stream


class Instances:
    """trees[i] placed with the row-vector 4x4 matrices world_to_object[i] /
    object_to_world[i] (sprout's Composed_transformation); the Tree objects
    are kept alive as long as this is."""

    def __init__(self, trees: list[Tree], world_to_object, object_to_world,
                 settings=(16, 64, 4), stream=None):
        w2o = np.ascontiguousarray(world_to_object, np.float32).reshape(-1, 16)
        o2w = np.ascontiguousarray(object_to_world, np.float32).reshape(-1, 16)
        if len(trees) != len(w2o) or len(trees) != len(o2w):
            raise ValueError("one tree and two matrices per instance")
        self._trees = list(trees)
        handles = np.array([t.handle for t in trees], np.uint32)
        st = _lib.MiBvhSettings(*settings)
        out = ctypes.c_uint32(MI_NULL)
        check(_lib.load().mi_instances_create(handles.ctypes.data, w2o.ctypes.data, o2w.ctypes.data,
                                              len(trees), ctypes.byref(st), _stream(stream),
                                              ctypes.byref(out)), "mi_instances_create")
        self.handle = out.value
        self.num_instances = len(trees)

    def release(self) -> None:
        if getattr(self, "handle", None) is not None:
            check(_lib.load().mi_instances_release(self.handle), "mi_instances_release")
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def intersect(self, rays: torch.Tensor, stream=None) -> torch.Tensor:
        """Closest hits of (n, 8) float32 device rays: (n, 8) int32 words
        {t, u, v, triangle, instance, 0, 0, 0}."""
        rays = _dev(rays, torch.float32, 8)
        n = rays.shape[0]
        hits = torch.empty((n, 8), dtype=torch.int32, device=rays.device)
        check(_lib.load().mi_trace_instances_closest(self.handle, rays.data_ptr(), n,
                                                     hits.data_ptr(), _stream(stream)),
              "mi_trace_instances_closest")
        return hits

    def intersect_p(self, rays: torch.Tensor, stream=None) -> torch.Tensor:
        rays = _dev(rays, torch.float32, 8)
        n = rays.shape[0]
        occ = torch.empty((n,), dtype=torch.int32, device=rays.device)
        check(_lib.load().mi_trace_instances_any(self.handle, rays.data_ptr(), n, occ.data_ptr(),
                                                 _stream(stream)), "mi_trace_instances_any")
        return occ

    def download(self):
        """(top-level nodes, instance per slot, world boxes (n, 6))"""
        nn, ns = ctypes.c_uint32(0), ctypes.c_uint32(0)
        lib = _lib.load()
        check(lib.mi_instances_download(self.handle, None, None, None, ctypes.byref(nn),
                                        ctypes.byref(ns)), "mi_instances_download")
        nodes = np.zeros(nn.value, NODE_DTYPE)
        slots = np.zeros(ns.value, np.uint32)
        boxes = np.zeros((self.num_instances, 6), np.float32)
        check(lib.mi_instances_download(self.handle, nodes.ctypes.data, slots.ctypes.data,
                                        boxes.ctypes.data, ctypes.byref(nn), ctypes.byref(ns)),
              "mi_instances_download")
        return nodes, slots, boxes


def random_transforms(n: int, seed: int, extent: float = 4.0, scale=(0.3, 1.2)):
    """n seeded random similarity transforms (rotation from a uniform random
    quaternion, uniform scale, position in [-extent, extent]^3) as row-vector
    4x4 matrices: (world_to_object, object_to_world), each (n, 16) float32.
    world_to_object is the inverse computed in float64 and rounded."""
    rng = np.random.default_rng(seed)
    o2w = np.zeros((n, 4, 4))
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x, y, z, w = q
        r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)],
                      [2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)],
                      [2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)]])
        s = rng.uniform(*scale)
        o2w[i, :3, :3] = s * r
        o2w[i, 3, :3] = rng.uniform(-extent, extent, 3)
        o2w[i, 3, 3] = 1.0
    w2o = np.linalg.inv(o2w)
    return (w2o.astype(np.float32).reshape(n, 16), o2w.astype(np.float32).reshape(n, 16))
