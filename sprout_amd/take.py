"""sprout's scene formats at the boundary, for the shapes and materials the
GPU path renders: triangle meshes in sprout's JSON mesh format
(triangle_json_handler.cpp: "geometry": "parts" {material_index,
start_index, num_indices}, "vertices": "positions", "indices"), scenes
(scene_loader.cpp: "materials" with Substitute {color, roughness, metallic,
ior} or Light {emission [r, g, b] | emittance {value, spectrum {rgb}}},
"entities" of type Prop / Light with a mesh file shape, materials and a
transformation {position, scale, rotation in degrees}) and takes
(take_loader.cpp: "scene", a Perspective camera {parameters.fov,
transformation, sensor.resolution, sensor.filter}, sampler
samples_per_pixel, integrator surface PTMIS min / max bounces).

Rotations follow json::create_rotation_matrix (base/json/json.cpp:216-227:
Rz Rx Ry from degrees) and transforms sprout's row-vector convention (a
point's world position is x r0 + y r1 + z r2 + r3 with rows r_i = scale_i
R_i).  sprout's camera looks along its +z with +x to the right
(camera_perspective.cpp:204-226) and its fov is horizontal; this library's
camera takes a vertical fov and puts fwd x up to the right, so a take's
image comes out mirrored left to right (`Take.mirror_x`): flip its columns.

Built-in shapes (Sphere, Rectangle, ...) and spectra given as temperatures
are not read: they raise NotImplementedError naming what is missing.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np

from .scenes import LAMBERT, LIGHT, MATERIAL_DTYPE, SUBSTITUTE


def _read(path):
    with open(path) as f:
        return json.load(f)


def load_mesh(path):
    """A triangle mesh file -> (positions (v, 3) float32, indices (t, 3)
    uint32, part of each triangle (t,) uint32)."""
    g = _read(path)["geometry"]
    pos = np.asarray(g["vertices"]["positions"], np.float32).reshape(-1, 3)
    idx = np.asarray(g["indices"], np.uint32).reshape(-1, 3)
    part = np.zeros(len(idx), np.uint32)
    for p in g.get("parts", [{"material_index": 0, "start_index": 0, "num_indices": 3 * len(idx)}]):
        first = int(p["start_index"]) // 3
        part[first:first + int(p["num_indices"]) // 3] = int(p["material_index"])
    if len(idx) and int(idx.max()) >= len(pos):
        raise ValueError(f"{path}: index {int(idx.max())} of {len(pos)} vertices")
    return pos, idx, part


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


def _material(node):
    rec = np.zeros(1, MATERIAL_DTYPE)[0]
    rec["ior"] = 1.46
    r = node.get("rendering", node)
    if "Substitute" in r:
        m = r["Substitute"]
        rec["kind"] = SUBSTITUTE
        rec["color"] = m.get("color", [0.5, 0.5, 0.5])
        rec["roughness"] = m.get("roughness", 0.8)
        rec["metallic"] = m.get("metallic", 0.0)
        rec["ior"] = m.get("ior", 1.46)
    elif "Lambert" in r:
        rec["kind"] = LAMBERT
        rec["color"] = r["Lambert"].get("color", [0.5, 0.5, 0.5])
    elif "Light" in r:
        m = r["Light"]
        rec["kind"] = LIGHT
        rec["ior"] = 1.0
        if "emission" in m:
            rec["emission"] = m["emission"]
        else:
            e = m.get("emittance", {})
            if e.get("quantity", "Radiance") != "Radiance":
                raise NotImplementedError(f"light emittance quantity {e.get('quantity')}")
            spec = e.get("spectrum", {})
            if "rgb" not in spec:
                raise NotImplementedError(f"light spectrum {sorted(spec)}")
            rec["emission"] = np.asarray(spec["rgb"], np.float32) * float(e.get("value", 1.0))
    else:
        raise NotImplementedError(f"material {sorted(r)}")
    return rec


def load_scene(path):
    """A scene file -> (positions, indices, per-triangle material, materials)
    of all its mesh entities in world space (transformed in float64, rounded
    to float32)."""
    doc = _read(path)
    base = os.path.dirname(os.path.abspath(path))
    names = {}
    mats = []
    for k, node in enumerate(doc.get("materials", [])):
        names[node.get("name", str(k))] = len(mats)
        mats.append(_material(node))
    pos, idx, tmat = [], [], []
    nv = 0
    for e in doc.get("entities", []):
        if e.get("type") not in ("Prop", "Light"):
            continue
        shape = e.get("shape", {})
        if "file" not in shape:
            raise NotImplementedError(f"shape {shape.get('type', shape)}")
        p, i, part = load_mesh(os.path.join(base, shape["file"]))
        slots = [names[m] for m in e.get("materials", [])]
        if not slots or int(part.max(initial=0)) >= len(slots):
            raise ValueError(f"entity {e.get('name', '?')}: {len(slots)} materials for its parts")
        m = object_to_world(e.get("transformation"))
        wp = (p.astype(np.float64) @ m[:3, :3] + m[3, :3]).astype(np.float32)
        pos.append(wp)
        idx.append(i + nv)
        tmat.append(np.asarray(slots, np.uint32)[part])
        nv += len(wp)
    if not idx:
        raise ValueError(f"{path}: no mesh entities")
    return (np.concatenate(pos), np.concatenate(idx).astype(np.uint32), np.concatenate(tmat),
            np.array(mats, MATERIAL_DTYPE))


class Take:
    """A take file: the scene arrays, the camera (for render.camera), the
    samples per pixel, the bounces and the sensor's filter."""

    def __init__(self, path):
        doc = _read(path)
        base = os.path.dirname(os.path.abspath(path))
        self.scene = load_scene(os.path.join(base, doc["scene"]))
        cam = doc.get("camera", {}).get("Perspective")
        if cam is None:
            raise NotImplementedError(f"camera {sorted(doc.get('camera', {}))}")
        sensor = cam.get("sensor", {})
        self.width, self.height = (int(v) for v in sensor.get("resolution", [1280, 720]))
        m = object_to_world(cam.get("transformation"))
        eye = m[3, :3]
        fov_h = math.radians(float(cam.get("parameters", {}).get("fov", 60)))
        fov_v = 2.0 * math.atan(math.tan(0.5 * fov_h) * self.height / self.width)
        self.camera = {"eye": tuple(eye), "target": tuple(eye + m[2, :3]), "up": tuple(m[1, :3]),
                       "fov_deg": math.degrees(fov_v)}
        self.mirror_x = True
        filt = sensor.get("filter", {})
        self.filter = next(iter(filt)) if filt else None
        self.filter_params = filt.get(self.filter, {}) if filt else {}
        spl = doc.get("sampler", {})
        self.samples_per_pixel = int(next(iter(spl.values())).get("samples_per_pixel", 1)) if spl else 1
        surf = doc.get("integrator", {}).get("surface", {})
        pt = next(iter(surf.values())) if surf else {}
        self.min_bounces = int(pt.get("min_bounces", 4))
        self.max_bounces = int(pt.get("max_bounces", 8))
