"""Renders the golden takes with the reference's own renderer and writes
tests/golden/sprout_renders.npz, the images the oracle and the GPU path
tracer are checked against in expectation (tests/test_sprout_render.py,
tests/test_sprout_render_gpu.py).

The renderer is the reference's core/, base/ and extension/ sources compiled
here by oracle/ref_render.mk into oracle/_ref/sprout_render (run from the
repository root, in the container that holds the reference):

    python tests/golden/make_sprout_renders.py

Each take is rendered as its file says, except: the sampler is Random with
SPP samples per pixel, and the sensor has no filter (each sample counts in
its own pixel, as the oracle's and the GPU's plain renders count it);
"small_filtered" keeps the take's Gaussian filter (radius 2, alpha 1.8).
The image is the resolved sensor before any postprocessor, in linear ACEScg.

It also writes take_smooth/sphere.json, the mesh of the second take: a UV
sphere with explicit unit normals in two parts (upper and lower half), and
take_room/, the box room of the path-tracer benchmark (sprout_amd.scenes.room)
in sprout's formats: its walls as one prop with a part per material (the
benchmark's Lambert walls, a material sprout lacks, as Substitutes of
roughness 1), its ceiling light as a Light entity, the benchmark's camera as a Perspective
camera (looking down -z: rotation y 180 degrees; the vertical fov of 40
degrees is also the horizontal one of a square image), 64 x 64.
"""

import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "sprout_render")
SPP = 65536
TAKES = {"small": os.path.join(HERE, "take_fixture", "small.take"),
         "smooth": os.path.join(HERE, "take_smooth", "smooth.take"),
         "room": os.path.join(HERE, "take_room", "room.take"),
         "small_filtered": os.path.join(HERE, "take_fixture", "small.take")}


def sphere_mesh(segments=12, rings=6, radius=0.7):
    """positions, normals, indices and parts of a UV sphere: the upper half
    is part 0, the lower half part 1"""
    pos, nrm = [], []
    for r in range(rings + 1):
        th = math.pi * r / rings
        for s in range(segments + 1):
            ph = 2 * math.pi * s / segments
            n = (math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph))
            nrm.append([round(c, 6) for c in n])
            pos.append([round(radius * c, 6) for c in n])
    upper, lower = [], []
    for r in range(rings):
        for s in range(segments):
            a = r * (segments + 1) + s
            b = a + segments + 1
            tris = []
            if r > 0:
                tris.append([a, a + 1, b])
            if r < rings - 1:
                tris.append([a + 1, b + 1, b])
            (upper if r < rings // 2 else lower).extend(tris)
    idx = upper + lower
    parts = [{"material_index": 0, "start_index": 0, "num_indices": 3 * len(upper)},
             {"material_index": 1, "start_index": 3 * len(upper), "num_indices": 3 * len(lower)}]
    return pos, nrm, idx, parts


def write_sphere():
    pos, nrm, idx, parts = sphere_mesh()
    doc = {"geometry": {"parts": parts, "primitive_topology": "triangle_list",
                        "vertices": {"positions": [c for p in pos for c in p],
                                     "normals": [c for n in nrm for c in n]},
                        "indices": [i for t in idx for i in t]}}
    with open(os.path.join(HERE, "take_smooth", "sphere.json"), "w") as f:
        json.dump(doc, f)


def write_room():
    sys.path.insert(0, ROOT)
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    out = os.path.join(HERE, "take_room")
    os.makedirs(out, exist_ok=True)
    light = mats["kind"][tmat] == scenes.LIGHT

    def mesh(tris, parts_of):
        used = np.unique(idx[tris])
        remap = np.full(len(pos), -1, np.int64)
        remap[used] = np.arange(len(used))
        order = np.argsort(parts_of, kind="stable")
        parts = []
        for m in np.unique(parts_of):
            first = int(np.searchsorted(parts_of[order], m))
            count = int((parts_of == m).sum())
            parts.append({"material_index": int(m), "start_index": 3 * first,
                          "num_indices": 3 * count})
        return {"geometry": {"parts": parts, "primitive_topology": "triangle_list",
                             "vertices": {"positions": [round(float(c), 6) for c in
                                                        pos[used].reshape(-1)]},
                             "indices": remap[idx[tris][order]].reshape(-1).tolist()}}

    walls = np.nonzero(~light)[0]
    kinds = tmat[walls]
    slots = sorted(set(kinds.tolist()))
    with open(os.path.join(out, "walls.json"), "w") as f:
        json.dump(mesh(walls, np.array([slots.index(k) for k in kinds])), f)
    with open(os.path.join(out, "light.json"), "w") as f:
        json.dump(mesh(np.nonzero(light)[0], np.zeros(int(light.sum()), np.int64)), f)
    materials = []
    for k, m in enumerate(mats):
        if m["kind"] == scenes.LIGHT:
            materials.append({"name": f"m{k}", "rendering": {"Light": {
                "emission": [float(c) for c in m["emission"]]}}})
        else:
            materials.append({"name": f"m{k}", "rendering": {"Substitute": {
                "color": [float(c) for c in m["color"]], "roughness": float(m["roughness"]),
                "metallic": float(m["metallic"]), "ior": float(m["ior"])}}})
    lamp = int(np.unique(tmat[light])[0])
    scene = {"materials": materials,
             "entities": [{"type": "Prop", "shape": {"file": "walls.json"},
                           "materials": [f"m{k}" for k in slots]},
                          {"type": "Light", "shape": {"file": "light.json"},
                           "materials": [f"m{lamp}"]}]}
    with open(os.path.join(out, "room.scene"), "w") as f:
        json.dump(scene, f, indent=1)
    cam = scenes.ROOM_CAMERA
    take = {"scene": "room.scene",
            "camera": {"Perspective": {
                "parameters": {"fov": cam["fov_deg"]},
                "transformation": {"position": list(cam["eye"]), "rotation": [0, 180, 0]},
                "sensor": {"resolution": [64, 64]}}},
            "sampler": {"Random": {"samples_per_pixel": 16}},
            "integrator": {"surface": {"PTMIS": {"min_bounces": 4, "max_bounces": 8}}}}
    with open(os.path.join(out, "room.take"), "w") as f:
        json.dump(take, f, indent=1)


def render(name, path):
    with open(path) as f:
        take = json.load(f)
    take["sampler"] = {"Random": {"samples_per_pixel": SPP}}
    sensor = take["camera"]["Perspective"].setdefault("sensor", {})
    if not name.endswith("_filtered"):
        sensor.pop("filter", None)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "image.f32")
        subprocess.run([BIN, os.path.dirname(path), json.dumps(take), str(os.cpu_count() or 1),
                        out], check=True, capture_output=True)
        raw = np.fromfile(out, np.float32)
    w, h = (int(v) for v in raw[:2].view(np.int32))
    img = raw[2:].reshape(h, w, 4)[..., :3].copy()
    print(f"{name}: {w}x{h}, {SPP} spp, mean {img.mean(axis=(0, 1))}")
    return img


def main() -> int:
    subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref_render.mk"), "-j8"], cwd=ROOT,
                   check=True)
    write_sphere()
    write_room()
    images = {name: render(name, path) for name, path in TAKES.items()}
    np.savez_compressed(os.path.join(HERE, "sprout_renders.npz"), spp=np.int64(SPP), **images)
    return 0


if __name__ == "__main__":
    sys.exit(main())
