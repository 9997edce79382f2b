"""Participating media of a scene: a density grid filling an axis-aligned
box, tracked by the path tracer (sprout's Tracking_multi with a volumetric
Grid material, core/rendering/integrator/volume/tracking_multi.cpp:44-183,
core/scene/material/volumetric/volumetric_grid.cpp:18-70)."""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

TRANSMITTANCE = 1  # MI_MEDIUM_TRANSMITTANCE
NEAREST = 2        # MI_MEDIUM_NEAREST

MAJORANTS = {"gridtree": 0, "flat": 1}  # MI_MAJORANTS_GRIDTREE, MI_MAJORANTS_FLAT


def _majorants(kind: str) -> int:
    if kind not in MAJORANTS:
        raise ValueError(f"majorants {kind!r} not one of {sorted(MAJORANTS)}")
    return MAJORANTS[kind]


def set_medium(scene, density, box_lo, box_hi, cc_a, cc_s, g: float,
               majorants: str = "gridtree") -> None:
    """density: (nz, ny, nx) float32 >= 0 (x fastest) stretched over the box
    [box_lo, box_hi]; cc_a / cc_s: absorption / scattering per unit density
    (rgb); g: Henyey-Greenstein anisotropy in (-1, 1); majorants: "gridtree"
    (sprout's adaptive octree of majorants) or "flat" (one majorant per
    8^3-voxel cell)."""
    kind = _majorants(majorants)
    d = np.ascontiguousarray(density, np.float32)
    if d.ndim != 3:
        raise ValueError("density must be (nz, ny, nx)")
    dims = np.array(d.shape[::-1], np.uint32)
    f3 = lambda v: np.ascontiguousarray(v, np.float32).reshape(3)  # noqa: E731
    lo, hi, a, s = f3(box_lo), f3(box_hi), f3(cc_a), f3(cc_s)
    check(_lib.load().mi_scene_set_medium_majorants(
        scene.handle, d.ctypes.data, dims.ctypes.data, lo.ctypes.data, hi.ctypes.data,
        a.ctypes.data, s.ctypes.data, float(g), kind), "mi_scene_set_medium_majorants")


def clear_medium(scene) -> None:
    check(_lib.load().mi_scene_set_medium(scene.handle, None, None, None, None, None, None, 0.0),
          "mi_scene_set_medium")


def gridtree(density, cc_a, cc_s, majorants: str = "gridtree"):
    """Host majorant gridtree of a density grid: (nodes uint32, data (n, 4)
    float32 {minorant mu_a, mu_s, majorant mu_a, mu_s}, num_cells (3,));
    majorants "flat" splits every cell to its 8^3-voxel leaves."""
    kind = _majorants(majorants)
    d = np.ascontiguousarray(density, np.float32)
    dims = np.array(d.shape[::-1], np.uint32)
    a = np.ascontiguousarray(cc_a, np.float32)
    s = np.ascontiguousarray(cc_s, np.float32)
    cells = np.zeros(3, np.uint32)
    nn, nd = ctypes.c_uint32(0), ctypes.c_uint32(0)
    lib = _lib.load()
    check(lib.mi_medium_majorant_tree(d.ctypes.data, dims.ctypes.data, a.ctypes.data,
                                      s.ctypes.data, kind, cells.ctypes.data, None,
                                      ctypes.byref(nn), None, ctypes.byref(nd)),
          "mi_medium_majorant_tree")
    nodes = np.zeros(nn.value, np.uint32)
    data = np.zeros((max(nd.value, 1), 4), np.float32)
    check(lib.mi_medium_majorant_tree(d.ctypes.data, dims.ctypes.data, a.ctypes.data,
                                      s.ctypes.data, kind, cells.ctypes.data, nodes.ctypes.data,
                                      ctypes.byref(nn), data.ctypes.data, ctypes.byref(nd)),
          "mi_medium_majorant_tree")
    return nodes, data[:nd.value], cells


def track(scene, rays, depth: int = 0, seed: int = 1, transmittance: bool = False,
          nearest: bool = False) -> np.ndarray:
    """Batch tracking of rays (n, 8: origin, min_t, direction, max_t) through
    the scene's medium on the GPU -> (n, 5): {event, t, weight rgb} or, with
    transmittance, {transmitted, 0, tr rgb}."""
    import torch

    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    n = r.shape[0]
    out = torch.zeros((n, 5), dtype=torch.float32, device="cuda")
    flags = (TRANSMITTANCE if transmittance else 0) | (NEAREST if nearest else 0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(_lib.load().mi_medium_track(scene.handle, ctypes.c_void_p(r.data_ptr()), depth, seed,
                                      flags, n, ctypes.c_void_p(out.data_ptr()), stream),
          "mi_medium_track")
    return out.cpu().numpy()
