"""sprout's scene formats at the boundary, for the shapes and materials the
GPU path renders: triangle meshes in sprout's JSON mesh format
(triangle_json_handler.cpp: "geometry": "parts" {material_index,
start_index, num_indices}, "vertices": "positions", "indices") or its
binary mesh format (.sub, triangle_mesh_provider.cpp:435), scenes
(scene_loader.cpp: "materials" with Substitute {color, roughness, metallic,
ior} or Light {emission | emittance {quantity Radiance, value, spectrum}},
"entities" of type Prop / Light with a mesh file shape, materials and a
transformation {position, scale, rotation in degrees}) and takes
(take_loader.cpp: "scene", a Perspective camera {parameters.fov,
transformation, sensor.resolution, sensor.filter}, sampler Random or
Golden_ratio with samples_per_pixel, integrator surface PTMIS min / max
bounces).

Colours are read as sprout's read_color reads them (material_provider.cpp:
1030-1066) and mapped from sRGB to ACEScg (AP1) with spectrum::sRGB_to_AP1
(base/spectrum/aces.hpp:14-18), as sprout's default build does (SU_ACESCG,
material_provider.cpp:1022-1028); the mapping's fused multiply-adds follow
the reference as compiled.  Rotations follow json::create_rotation_matrix
(base/json/json.cpp:216-227: Rz Rx Ry from degrees) and transforms sprout's
row-vector convention (a point's world position is x r0 + y r1 + z r2 + r3
with rows r_i = scale_i R_i).  The camera is sprout's Perspective: its
entity's position and rotation rows and the take's horizontal fov
(camera_perspective.cpp:47-80, 204-226).

Anything this path does not render raises NotImplementedError naming it:
built-in shapes (Sphere, Rectangle, ...), spectra given as temperatures,
textures, emissive or coated Substitutes and their other optional keys,
light quantities other than Radiance, and materials other than Substitute
and Light (sprout has no Lambert material).
"""

from __future__ import annotations

import json
import math
import os

import numpy as np

from .scenes import LIGHT, MATERIAL_DTYPE, SUBSTITUTE


def _read(path):
    with open(path) as f:
        return json.load(f)


SUB_MAGIC = b"SUB\0"
# serialize's Vertex: position, normal, tangent (packed float3), uv (float2),
# bitangent sign (uint8), padding: 48 bytes in a file's interleaved stream,
# 45 per vertex over its separate streams (shape_vertex.cpp:7-9)
_VERTEX_BYTES = 48
_VERTEX_UNPADDED = 3 * 4 + 3 * 4 + 3 * 4 + 2 * 4 + 1
_INDEX_TYPES = {"Int16": (np.int16, True), "UInt16": (np.uint16, False),
                "Int32": (np.int32, True), "UInt32": (np.uint32, False)}


def load_binary_mesh(path):
    """sprout's binary mesh (file::Type::SUB) as Provider::load_binary reads
    it (triangle_mesh_provider.cpp:435-666) and serialize::Exporter::write
    writes it (triangle_mesh_exporter.cpp): the magic "SUB\\0", a uint64 JSON
    size, the JSON header ("geometry": "parts", "vertices" {binary {offset,
    size}, num_vertices, layout}, "indices" {binary, num_indices, encoding}),
    then the binary block.  Positions come first in the vertex data (one
    48-byte vertex each when the layout's Bitangent_sign is in stream 0,
    else the separate position stream); indices are UInt16 / UInt32, or
    Int16 / Int32 deltas chained over the parts in order
    (fill_triangles_delta).  Returns positions, indices, the part of each
    triangle and the vertex normals (the normal stream)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != SUB_MAGIC:
        raise ValueError(f"{path}: not a binary mesh")
    json_size = int(np.frombuffer(data, np.uint64, 1, 4)[0])
    head = json.loads(data[12:12 + json_size].decode())
    g = head.get("geometry")
    if g is None:
        raise ValueError(f"{path}: no geometry node")
    parts = [(int(p["start_index"]), int(p["num_indices"]), int(p["material_index"]))
             for p in g.get("parts", [])]
    v = g.get("vertices", {})
    vb = v.get("binary", {})
    v_off, v_size = int(vb.get("offset", 0)), int(vb.get("size", 0))
    num_vertices = int(v.get("num_vertices", 0))
    interleaved = any(e.get("semantic_name") == "Bitangent_sign" and int(e.get("stream", 0)) == 0
                      for e in v.get("layout", []))
    ib = g.get("indices", {})
    i_off, i_size = int(ib.get("binary", {}).get("offset", 0)), int(ib.get("binary", {}).get("size", 0))
    dtype, delta = _INDEX_TYPES.get(ib.get("encoding"), (np.uint32, False))
    num_indices = int(ib.get("num_indices", 0))
    if 0 == num_vertices:
        # files written before the count was stored
        num_vertices = v_size // _VERTEX_BYTES
        if not interleaved:
            i_off = num_vertices * _VERTEX_UNPADDED
    if any(e.get("semantic_name") == "Tangent_space" for e in v.get("layout", [])):
        raise NotImplementedError(f"{path}: tangent space stored as quaternions")
    start = 12 + json_size
    if interleaved:
        rec = np.frombuffer(data, np.uint8, num_vertices * _VERTEX_BYTES, start + v_off)
        rec = rec.reshape(-1, _VERTEX_BYTES)
        pos = rec[:, :12].copy().view(np.float32).reshape(-1, 3)
        nrm = rec[:, 12:24].copy().view(np.float32).reshape(-1, 3)
    else:
        pos = np.frombuffer(data, np.float32, 3 * num_vertices, start + v_off).reshape(-1, 3).copy()
        # the normal stream follows the positions (Vertex_stream_separate)
        nrm = np.frombuffer(data, np.float32, 3 * num_vertices,
                            start + v_off + 12 * num_vertices).reshape(-1, 3).copy()
    if 0 == num_indices:
        num_indices = i_size // np.dtype(dtype).itemsize
    if start + i_off + i_size > len(data):
        raise ValueError(f"{path}: could not read all index data")
    raw = np.frombuffer(data, dtype, num_indices, start + i_off).astype(np.int64)
    idx = np.zeros(num_indices, np.int64)
    part = np.zeros(num_indices // 3, np.uint32)
    covered = np.zeros(num_indices // 3, bool)
    previous = 0
    for first, count, material in parts:
        if first + count > num_indices:
            raise ValueError(f"{path}: part indices out of bounds")
        t0, t1 = first // 3, (first + count) // 3
        seg = raw[3 * t0:3 * t1]
        if delta:
            seg = previous + np.cumsum(seg)
            if len(seg):
                previous = int(seg[-1])
        idx[3 * t0:3 * t1] = seg
        part[t0:t1] = material
        covered[t0:t1] = True
    if not covered.all():
        raise ValueError(f"{path}: triangles outside every part")
    if len(idx) and (idx.min() < 0 or idx.max() >= len(pos)):
        raise ValueError(f"{path}: index out of range of {len(pos)} vertices")
    return pos, idx.astype(np.uint32).reshape(-1, 3), part, nrm


def face_normals_last_writer(pos, idx):
    """The vertex normals sprout gives a JSON mesh without normals
    (triangle_mesh_provider.cpp:103-118): each triangle in order writes its
    normalized geometric normal, normalize(cross(b - a, c - a)) in float32,
    to its three vertices, so a vertex shared by several triangles keeps the
    last one's."""
    p = np.asarray(pos, np.float32)
    a, b, c = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    n = np.cross(b - a, c - a).astype(np.float32)
    n = (n / np.sqrt((n * n).sum(1, keepdims=True))).astype(np.float32)
    out = np.zeros_like(p)
    order = np.repeat(np.arange(len(idx)), 3)
    verts = idx.reshape(-1)
    last = np.full(len(p), -1, np.int64)
    np.maximum.at(last, verts, order)
    written = last >= 0
    out[written] = n[last[written]]
    return out


def load_mesh(path, with_normals=False):
    """A triangle mesh file, sprout's JSON mesh format or its binary format
    (load_binary_mesh) -> (positions (v, 3) float32, indices (t, 3) uint32,
    part of each triangle (t,) uint32), and with_normals: the vertex normals
    (v, 3) float32 sprout shades with (the file's, or for a JSON mesh
    without them face_normals_last_writer)."""
    with open(path, "rb") as f:
        binary = f.read(4) == SUB_MAGIC
    if binary:
        pos, idx, part, nrm = load_binary_mesh(path)
        return (pos, idx, part, nrm) if with_normals else (pos, idx, part)
    g = _read(path)["geometry"]
    pos = np.asarray(g["vertices"]["positions"], np.float32).reshape(-1, 3)
    idx = np.asarray(g["indices"], np.uint32).reshape(-1, 3)
    part = np.zeros(len(idx), np.uint32)
    for p in g.get("parts", [{"material_index": 0, "start_index": 0, "num_indices": 3 * len(idx)}]):
        first = int(p["start_index"]) // 3
        part[first:first + int(p["num_indices"]) // 3] = int(p["material_index"])
    if len(idx) and int(idx.max()) >= len(pos):
        raise ValueError(f"{path}: index {int(idx.max())} of {len(pos)} vertices")
    if not with_normals:
        return pos, idx, part
    if "normals" in g["vertices"]:
        nrm = np.asarray(g["vertices"]["normals"], np.float32).reshape(-1, 3)
        if nrm.shape != pos.shape:
            raise ValueError(f"{path}: {len(nrm)} normals for {len(pos)} vertices")
    else:
        nrm = face_normals_last_writer(pos, idx)
    return pos, idx, part, nrm


def rotation_matrix(degrees):
    """json::create_rotation_matrix: Rz Rx Ry (rows as matrix3x3.inl:61-103)"""
    ax, ay, az = (math.radians(float(d)) for d in degrees)
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ rx @ ry


def object_to_world(trafo):
    """(4, 4) float64 row-vector matrix of a transformation node"""
    t = trafo or {}
    r = rotation_matrix(t.get("rotation", [0, 0, 0]))
    s = np.asarray(t.get("scale", [1, 1, 1]), np.float64)
    m = np.eye(4)
    m[:3, :3] = s[:, None] * r
    m[3, :3] = np.asarray(t.get("position", [0, 0, 0]), np.float64)
    return m


# spectrum::sRGB_to_AP1 coefficients (base/spectrum/aces.hpp:14-18), rows = output channel
_SRGB_TO_AP1 = np.array([[0.61309732, 0.33952285, 0.04737928],
                         [0.07019422, 0.91635557, 0.01345259],
                         [0.02061560, 0.10956983, 0.86981512]], np.float32)


def _fma32(a, b, c) -> np.float32:
    """a * b + c rounded once to float32 (nearest, ties to even), exactly."""
    from fractions import Fraction

    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    f = np.float32(float(exact))
    best = None
    for cand in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(cand):
            continue
        err = abs(Fraction(float(cand)) - exact)
        key = (err, int(np.asarray(cand).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return best[1]


def srgb_to_ap1(rgb) -> np.ndarray:
    """spectrum::sRGB_to_AP1 in float32, as the reference build evaluates
    c0 x + c1 y + c2 z: fma(c2, z, fma(c0, x, c1 y))."""
    x, y, z = (np.float32(v) for v in rgb)
    out = np.zeros(3, np.float32)
    for k in range(3):
        c0, c1, c2 = _SRGB_TO_AP1[k]
        out[k] = _fma32(c2, z, _fma32(c0, x, np.float32(c1 * y)))
    return out


def _gamma_to_linear(c: np.float32) -> np.float32:
    """spectrum::gamma_to_linear_sRGB of one channel (rgb.hpp:34-44)"""
    c = np.float32(c)
    if c <= 0:
        return np.float32(0)
    if c < np.float32(0.04045):
        return np.float32(c / np.float32(12.92))
    if c < 1:
        return np.float32(np.power(np.float32((c + np.float32(0.055)) / np.float32(1.055)),
                                   np.float32(2.4)))
    return np.float32(1)


def read_color(value) -> np.ndarray:
    """read_color (material_provider.cpp:1030-1066), ACEScg build: a number
    is a grey whose mapped green channel fills all three; an array an sRGB
    colour; an object {"sRGB": colour, "linear": false} a colour read
    recursively (so mapped), then gamma-decoded as gamma_to_linear_sRGB(float3)
    does (its blue channel taken from green, rgb.hpp:54-57) and mapped again;
    a string "#rrggbb" a hex colour.  Temperatures are not read."""
    if isinstance(value, bool):
        raise NotImplementedError(f"colour {value!r}")
    if isinstance(value, (int, float)):
        g = srgb_to_ap1([value, value, value])[1]
        return np.array([g, g, g], np.float32)
    if isinstance(value, list):
        if len(value) != 3:
            raise ValueError(f"colour {value!r}")
        return srgb_to_ap1(value)
    if isinstance(value, dict):
        rgb = np.zeros(3, np.float32)
        linear = True
        for k, v in value.items():
            if k == "sRGB":
                rgb = read_color(v)
            elif k == "temperature":
                raise NotImplementedError("colour given as a black-body temperature")
            elif k == "linear":
                linear = bool(v)
        if not linear:
            rgb = np.array([_gamma_to_linear(rgb[0]), _gamma_to_linear(rgb[1]),
                            _gamma_to_linear(rgb[1])], np.float32)
        return srgb_to_ap1(rgb)
    if isinstance(value, str) and len(value) == 7 and value.startswith("#"):
        rgb = [np.float32(float(int(value[1 + 2 * i:3 + 2 * i], 16))) / np.float32(255)
               for i in range(3)]
        return srgb_to_ap1(rgb)
    raise NotImplementedError(f"colour {value!r}")


def _mapped_color(value, default) -> np.ndarray:
    """read_mapped_value for a colour (material_provider.cpp:1097-1111): an
    object is a texture description whose "value" is the colour (else the
    default stays); textures are not read"""
    if isinstance(value, dict):
        if value.get("file"):
            raise NotImplementedError(f"texture {value['file']}")
        if "value" in value:
            return read_color(value["value"])
        return np.asarray(default, np.float32)
    return read_color(value)


def _mapped_float(value, default) -> float:
    if isinstance(value, dict):
        if "file" in value:
            raise NotImplementedError(f"texture {value['file']}")
        return float(value.get("value", default))
    return float(value)


_SUBSTITUTE_KEYS = {"color", "roughness", "metallic", "ior", "sampler"}


def _material(node):
    rec = np.zeros(1, MATERIAL_DTYPE)[0]
    rec["ior"] = 1.46
    r = node.get("rendering", node)
    if "Substitute" in r:
        m = r["Substitute"]
        extra = sorted(k for k in m if k not in _SUBSTITUTE_KEYS and not (
            (k == "two_sided" and not m[k]) or (k == "anisotropy" and not m[k]) or
            (k == "emission" and not np.any(_mapped_color(m[k], np.zeros(3)))) or
            (k == "emission_factor")))
        if extra:
            raise NotImplementedError(f"Substitute keys {extra}")
        rec["kind"] = SUBSTITUTE
        rec["color"] = _mapped_color(m.get("color", {}), np.full(3, 0.5))
        rec["roughness"] = _mapped_float(m.get("roughness", 0.8), 0.8)
        rec["metallic"] = float(m.get("metallic", 0.0))
        rec["ior"] = float(m.get("ior", 1.46))
    elif "Light" in r:
        m = r["Light"]
        extra = sorted(k for k in m if k not in ("emission", "emittance", "sampler") and not (
            k == "two_sided" and not m[k]) and not (k == "emission_factor"))
        if extra:
            raise NotImplementedError(f"Light keys {extra}")
        rec["kind"] = LIGHT
        rec["ior"] = 1.0
        # Provider::load_light (material_provider.cpp:299-363): an emittance
        # of quantity Radiance gives value x colour; without a quantity the
        # emission (default 10, not mapped) is the radiance
        emission = _mapped_color(m.get("emission", {}), np.full(3, 10.0))
        quantity = ""
        color = np.ones(3, np.float32)
        value = np.float32(1)
        if "emittance" in m:
            e = m["emittance"]
            quantity = e.get("quantity", "")
            if "spectrum" in e:
                color = read_color(e["spectrum"])
            value = np.float32(e.get("value", 1.0))
        if quantity == "Radiance":
            rec["emission"] = (value * color).astype(np.float32)
        elif quantity == "":
            rec["emission"] = emission
        else:
            raise NotImplementedError(f"light emittance quantity {quantity}")
    else:
        raise NotImplementedError(f"material {sorted(r)} (the GPU path renders Substitute and "
                                  "Light; sprout has no Lambert material)")
    return rec


def load_scene(path, with_normals=False):
    """A scene file -> (positions, indices, per-triangle material, materials)
    of all its mesh entities in world space (transformed in float64, rounded
    to float32); with_normals: and the world shading normal of each vertex.

    sprout intersects a prop in object space and turns its geometric and
    shading normals to world space with the rotation alone
    (Mesh::intersect, triangle_mesh.cpp:388-421), so under a non-uniform
    scale its geometric normal is not the world triangle's.  The shading
    normals here are rotated alone as sprout's are; a prop whose scale tilts
    some triangle's geometric normal that way (by more than 1e-5) is refused,
    since this path takes the geometric normal from the world triangle."""
    doc = _read(path)
    base = os.path.dirname(os.path.abspath(path))
    names = {}
    mats = []
    for k, node in enumerate(doc.get("materials", [])):
        names[node.get("name", str(k))] = len(mats)
        mats.append(_material(node))
    pos, idx, tmat, nrms = [], [], [], []
    nv = 0
    for e in doc.get("entities", []):
        if e.get("type") not in ("Prop", "Light"):
            continue
        shape = e.get("shape", {})
        if "file" not in shape:
            raise NotImplementedError(f"shape {shape.get('type', shape)}")
        p, i, part, n = load_mesh(os.path.join(base, shape["file"]), with_normals=True)
        slots = [names[m] for m in e.get("materials", [])]
        if not slots or int(part.max(initial=0)) >= len(slots):
            raise ValueError(f"entity {e.get('name', '?')}: {len(slots)} materials for its parts")
        t = e.get("transformation") or {}
        m = object_to_world(t)
        wp = (p.astype(np.float64) @ m[:3, :3] + m[3, :3]).astype(np.float32)
        r = rotation_matrix(t.get("rotation", [0, 0, 0]))
        scale = np.asarray(t.get("scale", [1, 1, 1]), np.float64)
        if len(i) and not np.all(scale == scale[0]):
            _check_rotated_normals(p, i, wp, r, e.get("name", shape["file"]))
        wn = n.astype(np.float64) @ r
        length = np.sqrt((wn * wn).sum(1, keepdims=True))
        wn = np.where(length > 0, wn / np.where(length > 0, length, 1), 0)
        nrms.append(wn.astype(np.float32))
        pos.append(wp)
        idx.append(i + nv)
        tmat.append(np.asarray(slots, np.uint32)[part])
        nv += len(wp)
    if not idx:
        raise ValueError(f"{path}: no mesh entities")
    out = (np.concatenate(pos), np.concatenate(idx).astype(np.uint32), np.concatenate(tmat),
           np.array(mats, MATERIAL_DTYPE))
    return (*out, np.concatenate(nrms)) if with_normals else out


def _check_rotated_normals(p, i, wp, r, name):
    """refuses a prop whose geometric normals sprout would tilt (the object
    triangle's normal rotated alone against the world triangle's)"""
    def unit(v):
        return v / np.maximum(np.sqrt((v * v).sum(1, keepdims=True)), 1e-30)

    po = p.astype(np.float64)
    pw = wp.astype(np.float64)
    no = unit(np.cross(po[i[:, 1]] - po[i[:, 0]], po[i[:, 2]] - po[i[:, 0]])) @ r
    nw = unit(np.cross(pw[i[:, 1]] - pw[i[:, 0]], pw[i[:, 2]] - pw[i[:, 0]]))
    if np.abs(no - nw).max() > 1e-5:
        raise NotImplementedError(
            f"entity {name}: a non-uniform scale that tilts its normals (sprout rotates "
            "normals without the scale; bake the scale into the mesh)")


class Take:
    """A take file: the scene arrays, the vertex shading normals
    (render.Scene(..., normals=take.normals)), the camera (position,
    rotation rows and horizontal fov: render.perspective(width, height,
    **camera)), the sampler, the samples per pixel, the bounces and the
    sensor's filter."""

    def __init__(self, path):
        doc = _read(path)
        base = os.path.dirname(os.path.abspath(path))
        *self.scene, self.normals = load_scene(os.path.join(base, doc["scene"]), with_normals=True)
        self.scene = tuple(self.scene)
        cam = doc.get("camera", {}).get("Perspective")
        if cam is None:
            raise NotImplementedError(f"camera {sorted(doc.get('camera', {}))}")
        params = cam.get("parameters", {})
        unsupported = sorted(k for k in params if k not in ("fov",) and not (
            k == "lens_radius" and not params[k]))
        if unsupported:
            raise NotImplementedError(f"camera parameters {unsupported}")
        sensor = cam.get("sensor", {})
        self.width, self.height = (int(v) for v in sensor.get("resolution", [1280, 720]))
        m = object_to_world(cam.get("transformation"))
        self.camera = {"position": m[3, :3].astype(np.float32),
                       "rotation": m[:3, :3].reshape(9).astype(np.float32),
                       "fov_deg": float(params.get("fov", 60))}
        filt = sensor.get("filter", {})
        self.filter = next(iter(filt)) if filt else None
        self.filter_params = filt.get(self.filter, {}) if filt else {}
        spl = doc.get("sampler", {})
        self.sampler = next(iter(spl)) if spl else "Random"
        if self.sampler not in ("Random", "Golden_ratio"):
            raise NotImplementedError(f"sampler {self.sampler}")
        self.samples_per_pixel = int(spl[self.sampler].get("samples_per_pixel", 1)) if spl else 1
        surf = doc.get("integrator", {}).get("surface", {})
        pt = next(iter(surf.values())) if surf else {}
        self.min_bounces = int(pt.get("min_bounces", 4))
        self.max_bounces = int(pt.get("max_bounces", 8))

    @property
    def golden_ratio(self) -> bool:
        """the take's sampler is Golden_ratio (render(..., golden_ratio=True))"""
        return self.sampler == "Golden_ratio"
