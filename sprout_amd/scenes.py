"""Synthetic inputs of the BVH micro-benchmark (host-side generators in the C ABI).

Mesh: an icosphere subdivided `subdivisions` times (20 * 4^s triangles),
radially displaced by `octaves`-octave value noise of amplitude `amplitude`,
plus `num_random` small triangles in the unit cube whose vertices sit at
log-normal(mu, sigma) distances from a random centre.  At subdivisions=8,
octaves=4, amplitude=0.1, num_random=100000, mu=-5, sigma=0.5 this is the
1.41 M-triangle micro-benchmark mesh.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

FLT_MAX = np.float32(3.4028234663852886e38)


def displaced_icosphere(subdivisions=8, octaves=4, amplitude=0.1, noise_seed=1, num_random=100000,
                        log_mu=-5.0, log_sigma=0.5, random_seed=2):
    lib = _lib.load()
    nv = ctypes.c_uint32(0)
    nt = ctypes.c_uint32(0)
    check(lib.mi_gen_displaced_icosphere(subdivisions, octaves, amplitude, noise_seed, num_random,
                                         log_mu, log_sigma, random_seed, None, ctypes.byref(nv),
                                         None, ctypes.byref(nt)), "mi_gen_displaced_icosphere")
    pos = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    check(lib.mi_gen_displaced_icosphere(subdivisions, octaves, amplitude, noise_seed, num_random,
                                         log_mu, log_sigma, random_seed, pos.ctypes.data,
                                         ctypes.byref(nv), idx.ctypes.data, ctypes.byref(nt)),
          "mi_gen_displaced_icosphere")
    return pos, idx


def camera_rays(width=2048, height=2048, eye=(0.6, 0.9, 3.2), target=(0.0, 0.0, 0.0),
                up=(0.0, 1.0, 0.0), fov_deg=45.0) -> np.ndarray:
    """(width*height, 8) float32 pinhole rays, min_t 0, max_t FLT_MAX."""
    rays = np.zeros((width * height, 8), np.float32)
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    check(_lib.load().mi_gen_camera_rays(width, height, e.ctypes.data, t.ctypes.data, u.ctypes.data,
                                         fov_deg, rays.ctypes.data), "mi_gen_camera_rays")
    return rays


def bounce_rays(rays: np.ndarray, hits: np.ndarray, positions: np.ndarray, indices: np.ndarray,
                num_out: int, seed: int = 3) -> np.ndarray:
    """Cosine bounce rays from the closest hits (hits as (n, 4) 32-bit words)."""
    rays = np.ascontiguousarray(rays, np.float32)
    hits = np.ascontiguousarray(hits).view(np.uint32)
    positions = np.ascontiguousarray(positions, np.float32)
    indices = np.ascontiguousarray(indices, np.uint32)
    out = np.zeros((num_out, 8), np.float32)
    check(_lib.load().mi_gen_bounce_rays(rays.ctypes.data, hits.ctypes.data, len(rays),
                                         positions.ctypes.data, 3, indices.ctypes.data, seed,
                                         out.ctypes.data, num_out), "mi_gen_bounce_rays")
    return out


# mi_material (include/sprout_gpu.h), 40 bytes
MATERIAL_DTYPE = np.dtype([("color", "<f4", 3), ("emission", "<f4", 3), ("roughness", "<f4"),
                           ("metallic", "<f4"), ("ior", "<f4"), ("kind", "<u4")])

SUBSTITUTE, LAMBERT, LIGHT = 0, 1, 2

# The room seen through its open +z side.
ROOM_CAMERA = {"eye": (0.0, 1.0, 3.4), "target": (0.0, 1.0, 0.0), "up": (0.0, 1.0, 0.0),
               "fov_deg": 40.0}


def room(slit: float = 0.0):
    """The procedural box room: 36 triangles (46 with the slit enclosure).

    Returns positions (v, 3) f32, indices (t, 3) u32, per-triangle material
    (t,) u32 and materials (m,) MATERIAL_DTYPE: 0 grey 0.75, 1 red
    0.63/0.06/0.05, 2 green 0.14/0.45/0.09 (all Lambert), 3 the light of
    radiance 17/12/4."""
    lib = _lib.load()
    nv, nt, nm = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    check(lib.mi_gen_room(slit, None, ctypes.byref(nv), None, ctypes.byref(nt), None, None,
                          ctypes.byref(nm)), "mi_gen_room")
    pos = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    tmat = np.zeros(nt.value, np.uint32)
    mats = np.zeros(nm.value, MATERIAL_DTYPE)
    check(lib.mi_gen_room(slit, pos.ctypes.data, ctypes.byref(nv), idx.ctypes.data,
                          ctypes.byref(nt), tmat.ctypes.data, mats.ctypes.data, ctypes.byref(nm)),
          "mi_gen_room")
    return pos, idx, tmat, mats


# camera of the cloud scene: looking at the medium box from the front
CLOUD_CAMERA = {"eye": (0.0, 1.4, 4.6), "target": (0.0, 1.2, 0.0), "up": (0.0, 1.0, 0.0),
                "fov_deg": 40.0}

# the medium box of the cloud scene and its coefficients per unit density:
# scattering albedo 0.9, Henyey-Greenstein g 0.6
CLOUD_BOX = ((-1.0, 0.2, -1.0), (1.0, 2.2, 1.0))
CLOUD_CC_A = (0.1, 0.1, 0.1)
CLOUD_CC_S = (0.9, 0.9, 0.9)
CLOUD_G = 0.6


def cloud_scene():
    """Ground quad and a 1000-triangle emissive quad above it (1002
    triangles): positions, indices, per-triangle material, materials
    (0 ground Lambert 0.5, 1 light 4 / 4.5 / 5)."""
    lib = _lib.load()
    nv, nt, nm = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
    check(lib.mi_gen_cloud_scene(None, ctypes.byref(nv), None, ctypes.byref(nt), None, None,
                                 ctypes.byref(nm)), "mi_gen_cloud_scene")
    pos = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    tmat = np.zeros(nt.value, np.uint32)
    mats = np.zeros(nm.value, MATERIAL_DTYPE)
    check(lib.mi_gen_cloud_scene(pos.ctypes.data, ctypes.byref(nv), idx.ctypes.data,
                                 ctypes.byref(nt), tmat.ctypes.data, mats.ctypes.data,
                                 ctypes.byref(nm)), "mi_gen_cloud_scene")
    return pos, idx, tmat, mats


def fbm_density(n: int = 256, octaves: int = 6, seed: int = 1, max_density: float = 20.0):
    """(n, n, n) float32 density grid (z, y, x; x fastest) of seeded fBm in a
    spherical envelope, in [0, max_density]."""
    out = np.zeros((n, n, n), np.float32)
    check(_lib.load().mi_gen_fbm_density(n, n, n, octaves, seed, max_density, out.ctypes.data),
          "mi_gen_fbm_density")
    return out


# camera of the batched path-tracing scene
BATCH_CAMERA = {"eye": (0.0, 6.0, 26.0), "target": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
                "fov_deg": 45.0}


def batch_scene_instanced(num_instances: int = 64, num_lights: int = 10000, seed: int = 1):
    """The batched path-tracing workload as sprout's props: num_instances
    placements (seeded similarity transforms in a 20-unit cube) of 4 distinct
    20 480-triangle displaced spheres, each with its own Substitute material
    (roughness U[0.05, 1], metallic Bernoulli(0.3), base colour U[0.1, 0.9]^3);
    a grey ground plane under them; and num_lights small emissive triangles
    above them with log-normal (sigma 1) radiance.  Returns a dict: meshes
    [(positions, indices)], mesh_of, world_to_object, object_to_world,
    instance_materials, and the world triangles positions, indices,
    triangle_materials with the materials of all."""
    from .instances import random_transforms

    rng = np.random.default_rng(seed)
    meshes = [displaced_icosphere(5, octaves=4, amplitude=0.15, noise_seed=seed + 10 + k,
                                  num_random=0) for k in range(4)]
    w2o, o2w = random_transforms(num_instances, seed + 100, extent=10.0)
    mats = np.zeros(num_instances + 1, MATERIAL_DTYPE)
    for i in range(num_instances):
        mats[i]["kind"] = SUBSTITUTE
        mats[i]["color"] = rng.uniform(0.1, 0.9, 3)
        mats[i]["roughness"] = rng.uniform(0.05, 1.0)
        mats[i]["metallic"] = 1.0 if rng.random() < 0.3 else 0.0
        mats[i]["ior"] = 1.46
    # ground
    g = num_instances
    mats[g]["kind"] = LAMBERT
    mats[g]["color"] = (0.5, 0.5, 0.5)
    mats[g]["ior"] = 1.46
    gp = np.array([[-40, -12, -40], [40, -12, -40], [40, -12, 40], [-40, -12, 40]], np.float32)
    gi = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    # lights: small triangles facing down at y in [12, 16], one material each
    lm = np.zeros(num_lights, MATERIAL_DTYPE)
    lm["kind"] = LIGHT
    lm["ior"] = 1.0
    lm["emission"] = (20.0 * np.exp(rng.normal(size=(num_lights, 1)))
                      * rng.uniform(0.5, 1.0, (num_lights, 3)))
    c = np.stack([rng.uniform(-20, 20, num_lights), rng.uniform(12, 16, num_lights),
                  rng.uniform(-20, 20, num_lights)], axis=1)
    sz = rng.uniform(0.1, 0.4, (num_lights, 1))
    lp = np.stack([c, c + sz * [1, 0, 0], c + sz * [0, 0, 1]], axis=1).astype(np.float32)
    li = 4 + np.arange(3 * num_lights, dtype=np.uint32).reshape(-1, 3)
    # cross(b - a, c - a) = (1,0,0) x (0,0,1) points down
    return {
        "meshes": meshes,
        "mesh_of": (np.arange(num_instances) % 4).astype(np.uint32),
        "world_to_object": w2o,
        "object_to_world": o2w,
        "instance_materials": np.arange(num_instances, dtype=np.uint32),
        "positions": np.concatenate([gp, lp.reshape(-1, 3)]).astype(np.float32),
        "indices": np.concatenate([gi, li]).astype(np.uint32),
        "triangle_materials": np.concatenate(
            [np.full(2, g, np.uint32), np.arange(num_lights, dtype=np.uint32) + num_instances + 1]),
        "materials": np.concatenate([mats, lm]),
    }


def batch_scene(num_instances: int = 64, num_lights: int = 10000, seed: int = 1):
    """batch_scene_instanced baked into one triangle soup: every instance's
    mesh transformed to world space (in double, rounded to float), then the
    ground and the lights.  Returns positions, indices, per-triangle material
    and materials."""
    b = batch_scene_instanced(num_instances, num_lights, seed)
    pos, idx, tmat = [], [], []
    base = 0
    for i in range(num_instances):
        p, t = b["meshes"][b["mesh_of"][i]]
        m = b["object_to_world"][i].reshape(4, 4).astype(np.float64)
        wp = (p.astype(np.float64) @ m[:3, :3] + m[3, :3]).astype(np.float32)
        pos.append(wp)
        idx.append(t + base)
        tmat.append(np.full(len(t), b["instance_materials"][i], np.uint32))
        base += len(wp)
    pos.append(b["positions"])
    idx.append(b["indices"] + base)
    tmat.append(b["triangle_materials"])
    return (np.concatenate(pos).astype(np.float32), np.concatenate(idx).astype(np.uint32),
            np.concatenate(tmat).astype(np.uint32), b["materials"])
