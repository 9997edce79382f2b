"""Host side of the light tree: building one from light records
(mi_light_tree_build, no GPU needed), reading a scene's tree back, and batch
sampling / pdf queries on the device (mi_light_tree_sample / _pdf)."""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

MAX_LIGHTS = 64

# mi_light_node (48 bytes)
NODE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("cone", "<f4", 4),
                       ("power", "<f4"), ("variance", "<f4"), ("bits", "<u4"),
                       ("num_lights", "<u4")])
# per-light record of a scene (48 bytes)
LIGHT_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("cone", "<f4", 4),
                        ("power", "<f4"), ("two_sided", "<u4"), ("triangle", "<u4"),
                        ("area", "<f4")])


def build_tree(box_min, box_max, cones, powers, two_sided):
    """Tree of n lights -> (nodes, middles, mapping, orders)."""
    bmin = np.ascontiguousarray(box_min, np.float32)
    bmax = np.ascontiguousarray(box_max, np.float32)
    cn = np.ascontiguousarray(cones, np.float32)
    pw = np.ascontiguousarray(powers, np.float32)
    ts = np.ascontiguousarray(two_sided, np.uint8)
    n = len(pw)
    nn = ctypes.c_uint32(0)
    lib = _lib.load()
    check(lib.mi_light_tree_build(bmin.ctypes.data, bmax.ctypes.data, cn.ctypes.data,
                                  pw.ctypes.data, ts.ctypes.data, n, None, None, None, None,
                                  ctypes.byref(nn)), "mi_light_tree_build")
    nodes = np.zeros(max(nn.value, 1), NODE_DTYPE)
    middles = np.zeros(max(nn.value, 1), np.uint32)
    mapping = np.zeros(max(n, 1), np.uint32)
    orders = np.zeros(max(n, 1), np.uint32)
    check(lib.mi_light_tree_build(bmin.ctypes.data, bmax.ctypes.data, cn.ctypes.data,
                                  pw.ctypes.data, ts.ctypes.data, n, nodes.ctypes.data,
                                  middles.ctypes.data, mapping.ctypes.data, orders.ctypes.data,
                                  ctypes.byref(nn)), "mi_light_tree_build")
    return nodes[:nn.value], middles[:nn.value], mapping[:n], orders[:n]


def build_part_tree(box_min, box_max, cones, powers, part_cone):
    """Primitive_tree of one part's n triangles -> (nodes, middles, mapping,
    orders, cdf_offset, cdfs)."""
    bmin = np.ascontiguousarray(box_min, np.float32)
    bmax = np.ascontiguousarray(box_max, np.float32)
    cn = np.ascontiguousarray(cones, np.float32)
    pw = np.ascontiguousarray(powers, np.float32)
    pc = np.ascontiguousarray(part_cone, np.float32)
    n = len(pw)
    nn, nc = ctypes.c_uint32(0), ctypes.c_uint32(0)
    lib = _lib.load()
    check(lib.mi_light_part_tree_build(bmin.ctypes.data, bmax.ctypes.data, cn.ctypes.data,
                                       pw.ctypes.data, n, pc.ctypes.data, None, None, None, None,
                                       None, None, ctypes.byref(nn), ctypes.byref(nc)),
          "mi_light_part_tree_build")
    nodes = np.zeros(max(nn.value, 1), NODE_DTYPE)
    middles = np.zeros(max(nn.value, 1), np.uint32)
    mapping = np.zeros(max(n, 1), np.uint32)
    orders = np.zeros(max(n, 1), np.uint32)
    offs = np.zeros(max(nn.value, 1), np.uint32)
    cdfs = np.zeros(max(nc.value, 1), np.float32)
    check(lib.mi_light_part_tree_build(bmin.ctypes.data, bmax.ctypes.data, cn.ctypes.data,
                                       pw.ctypes.data, n, pc.ctypes.data, nodes.ctypes.data,
                                       middles.ctypes.data, mapping.ctypes.data, orders.ctypes.data,
                                       offs.ctypes.data, cdfs.ctypes.data, ctypes.byref(nn),
                                       ctypes.byref(nc)), "mi_light_part_tree_build")
    return (nodes[:nn.value], middles[:nn.value], mapping[:n], orders[:n], offs[:nn.value],
            cdfs[:nc.value])


def scene_tree(scene):
    """(nodes, middles, mapping, orders, lights) of a render.Scene."""
    lib = _lib.load()
    nn, nl = ctypes.c_uint32(0), ctypes.c_uint32(0)
    check(lib.mi_scene_light_tree(scene.handle, None, None, None, None, None, ctypes.byref(nn),
                                  ctypes.byref(nl)), "mi_scene_light_tree")
    nodes = np.zeros(max(nn.value, 1), NODE_DTYPE)
    middles = np.zeros(max(nn.value, 1), np.uint32)
    mapping = np.zeros(max(nl.value, 1), np.uint32)
    orders = np.zeros(max(nl.value, 1), np.uint32)
    lights = np.zeros(max(nl.value, 1), LIGHT_DTYPE)
    check(lib.mi_scene_light_tree(scene.handle, nodes.ctypes.data, middles.ctypes.data,
                                  mapping.ctypes.data, orders.ctypes.data, lights.ctypes.data,
                                  ctypes.byref(nn), ctypes.byref(nl)), "mi_scene_light_tree")
    return nodes[:nn.value], middles[:nn.value], mapping[:nl.value], orders[:nl.value], \
        lights[:nl.value]


def sample(scene, points, normals, randoms, split: bool, total_sphere: bool = False):
    """Device batch Tree::random_light -> (picks (n, 64) light ids, pdfs (n, 64), counts).
    total_sphere: the points lie in a medium (no horizon test)."""
    import torch

    p = torch.from_numpy(np.ascontiguousarray(points, np.float32)).cuda()
    nv = torch.from_numpy(np.ascontiguousarray(normals, np.float32)).cuda()
    r = torch.from_numpy(np.ascontiguousarray(randoms, np.float32)).cuda()
    n = len(r)
    picks = torch.zeros((n, MAX_LIGHTS, 2), dtype=torch.int32, device="cuda")
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(_lib.load().mi_light_tree_sample(scene.handle, ctypes.c_void_p(p.data_ptr()),
                                           ctypes.c_void_p(nv.data_ptr()),
                                           ctypes.c_void_p(r.data_ptr()),
                                           _flags(split, total_sphere), n,
                                           ctypes.c_void_p(picks.data_ptr()),
                                           ctypes.c_void_p(counts.data_ptr()), stream),
          "mi_light_tree_sample")
    pk = picks.cpu().numpy()
    return pk[:, :, 0].view(np.uint32), pk[:, :, 1].view(np.float32), \
        counts.cpu().numpy().astype(np.int64)


def _flags(split: bool, total_sphere: bool) -> int:
    return (1 if split else 0) | (2 if total_sphere else 0)  # MI_LIGHT_SPLIT, MI_LIGHT_TOTAL_SPHERE


def pdf(scene, points, normals, lights, split: bool, total_sphere: bool = False) -> np.ndarray:
    import torch

    p = torch.from_numpy(np.ascontiguousarray(points, np.float32)).cuda()
    nv = torch.from_numpy(np.ascontiguousarray(normals, np.float32)).cuda()
    li = torch.from_numpy(np.ascontiguousarray(lights, np.uint32).view(np.int32)).cuda()
    n = len(li)
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(_lib.load().mi_light_tree_pdf(scene.handle, ctypes.c_void_p(p.data_ptr()),
                                        ctypes.c_void_p(nv.data_ptr()),
                                        ctypes.c_void_p(li.data_ptr()),
                                        _flags(split, total_sphere), n,
                                        ctypes.c_void_p(out.data_ptr()), stream),
          "mi_light_tree_pdf")
    return out.cpu().numpy()
