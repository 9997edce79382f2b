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


def _device(a, dtype):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def build_tree_device(box_min, box_max, cones, powers, two_sided):
    """As build_tree, built on the device (mi_light_tree_build_device)."""
    import torch

    n = len(powers)
    args = [_device(box_min, np.float32), _device(box_max, np.float32), _device(cones, np.float32),
            _device(powers, np.float32), _device(two_sided, np.uint8)]
    nodes = torch.zeros((max(2 * n - 1, 1), NODE_DTYPE.itemsize // 4), dtype=torch.int32,
                        device="cuda")
    middles = torch.zeros(max(2 * n - 1, 1), dtype=torch.int32, device="cuda")
    mapping = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    orders = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    nn = ctypes.c_uint32(0)
    check(_lib.load().mi_light_tree_build_device(*[_dptr(a) for a in args], n, _dptr(nodes),
                                                 _dptr(middles), _dptr(mapping), _dptr(orders),
                                                 ctypes.byref(nn), _stream()),
          "mi_light_tree_build_device")
    torch.cuda.synchronize()
    k = nn.value
    return (nodes.cpu().numpy().view(NODE_DTYPE).reshape(-1)[:k],
            middles.cpu().numpy().view(np.uint32)[:k], mapping.cpu().numpy().view(np.uint32)[:n],
            orders.cpu().numpy().view(np.uint32)[:n])


def build_part_trees_device(box_min, box_max, cones, powers, first, part_cones):
    """The Primitive_trees of parts first[r] .. first[r + 1] - 1 at once on the
    device -> per part (nodes, middles, mapping, orders, cdf_offset, cdfs) as
    build_part_tree gives them."""
    import torch

    n = len(powers)
    fst = np.ascontiguousarray(first, np.uint32)
    r = len(fst) - 1
    args = [_device(box_min, np.float32), _device(box_max, np.float32), _device(cones, np.float32),
            _device(powers, np.float32)]
    pc = _device(part_cones, np.float32)
    words = NODE_DTYPE.itemsize // 4
    nodes = torch.zeros((2 * n, words), dtype=torch.int32, device="cuda")
    middles = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    mapping = torch.zeros(n, dtype=torch.int32, device="cuda")
    orders = torch.zeros(n, dtype=torch.int32, device="cuda")
    offs = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    cdfs = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    ncdf = torch.zeros(r, dtype=torch.int32, device="cuda")
    nn = np.zeros(r, np.uint32)
    check(_lib.load().mi_light_part_trees_build_device(
        *[_dptr(a) for a in args], n, fst.ctypes.data, r, _dptr(pc), _dptr(nodes), _dptr(middles),
        _dptr(mapping), _dptr(orders), _dptr(offs), _dptr(cdfs), nn.ctypes.data, _dptr(ncdf),
        _stream()), "mi_light_part_trees_build_device")
    torch.cuda.synchronize()
    nodes = nodes.cpu().numpy().view(NODE_DTYPE).reshape(-1)
    middles = middles.cpu().numpy().view(np.uint32)
    mapping = mapping.cpu().numpy().view(np.uint32)
    orders = orders.cpu().numpy().view(np.uint32)
    offs = offs.cpu().numpy().view(np.uint32)
    cdfs = cdfs.cpu().numpy()
    ncdf = ncdf.cpu().numpy()
    out = []
    for k in range(r):
        b, e = int(fst[k]), int(fst[k + 1])
        nb, cb = 2 * b - k, 2 * b
        m = int(nn[k])
        out.append((nodes[nb:nb + m], middles[nb:nb + m], mapping[b:e], orders[b:e],
                    offs[nb:nb + m], cdfs[cb:cb + int(ncdf[k])]))
    return out


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
