"""Writes tests/golden/mesh_parts.sub with the reference's own binary mesh
writer (serialize::Exporter::write, compiled into oracle/_ref by
oracle/ref.mk) from tests/golden/mesh_parts.json, a small two-part mesh in
sprout's JSON mesh format with normals and texture coordinates.

    python tests/golden/make_mesh_sub.py

The functions here also serve the CPU test that exports meshes of every
index encoding at test time when oracle/_ref is built.
"""

from __future__ import annotations

import ctypes
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def mesh_json(pos, idx, parts, normals=None, uvs=None) -> str:
    """sprout's JSON mesh text: parts [(start_index, num_indices, material)]."""
    v = {"positions": [float(x) for x in np.asarray(pos, np.float32).reshape(-1)]}
    if normals is not None:
        v["normals"] = [float(x) for x in np.asarray(normals, np.float32).reshape(-1)]
    if uvs is not None:
        v["texture_coordinates_0"] = [float(x) for x in np.asarray(uvs, np.float32).reshape(-1)]
    g = {"parts": [{"material_index": m, "start_index": s, "num_indices": n} for s, n, m in parts],
         "vertices": v, "indices": [int(i) for i in np.asarray(idx).reshape(-1)]}
    return json.dumps({"geometry": g})


def export(lib_path: str, text: str, directory: str, name: str) -> str:
    """<directory>/<name>.sub written by the reference's exporter."""
    lib = ctypes.CDLL(lib_path)
    lib.sref_export_mesh.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    cwd = os.getcwd()
    os.chdir(directory)  # the exporter writes into the current directory
    try:
        st = lib.sref_export_mesh(text.encode(), name.encode())
    finally:
        os.chdir(cwd)
    if st:
        raise RuntimeError("the reference's JSON handler rejected the mesh")
    return os.path.join(directory, name + ".sub")


def parts_mesh():
    """A fan of 7 triangles over 9 vertices, parts of 4 and 3 triangles."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(-1, 1, (9, 3)).astype(np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 5], [0, 5, 6], [0, 6, 7], [0, 7, 8]])
    normals = np.tile([0.0, 1.0, 0.0], (9, 1))
    uvs = rng.uniform(0, 1, (9, 2))
    return pos, idx, [(0, 12, 1), (12, 9, 0)], normals, uvs


def main() -> int:
    sys.path.insert(0, ROOT)
    from oracle import oracle as O

    lib = O.ref_lib_path()
    if not os.path.exists(lib):
        print("build oracle/_ref first: make -f oracle/ref.mk", file=sys.stderr)
        return 1
    pos, idx, parts, normals, uvs = parts_mesh()
    text = mesh_json(pos, idx, parts, normals, uvs)
    with open(os.path.join(HERE, "mesh_parts.json"), "w") as f:
        f.write(text)
    tmp = tempfile.mkdtemp()
    try:
        out = export(lib, text, tmp, "mesh_parts")
        shutil.copy(out, os.path.join(HERE, "mesh_parts.sub"))
    finally:
        shutil.rmtree(tmp)
    print("wrote", os.path.join(HERE, "mesh_parts.sub"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
