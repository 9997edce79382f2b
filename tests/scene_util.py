"""Helpers shared by the tests: the oracle's light tree of a scene.

The lights (one per emissive triangle) are computed by the oracle from the
geometry; the tree over them is built by the library's host builder
(mi_light_tree_build), whose splits the light-tree tests check separately.
"""

from oracle import oracle as O


def light_tree(positions, indices, triangle_materials, materials) -> O.LightTree:
    from sprout_amd import lights

    return O.scene_light_tree(positions, indices, triangle_materials, materials, lights.build_tree)


def room_with_lights(num_lights=48, seed=3, radiance_scale=20.0):
    """The box room plus num_lights small emissive triangles under the
    ceiling, facing down, with log-normal radiance (sigma 1): a scene whose
    light tree has several levels."""
    import numpy as np

    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    rng = np.random.default_rng(seed)
    new_pos, new_idx, new_tm = [pos], [idx], [tmat]
    new_mats = [mats]
    base = len(pos)
    for k in range(num_lights):
        c = np.array([rng.uniform(-0.9, 0.9), rng.uniform(1.5, 1.95), rng.uniform(-0.9, 0.9)])
        s = rng.uniform(0.02, 0.08)
        # wound so that cross(b - a, c - a) points down (-y)
        tri = np.array([c + [0, 0, 0], c + [s, 0, 0], c + [0, 0, s]], np.float32)
        new_pos.append(tri)
        new_idx.append(np.array([[base, base + 1, base + 2]], np.uint32))
        base += 3
        m = np.zeros(1, scenes.MATERIAL_DTYPE)
        m["kind"] = scenes.LIGHT
        m["emission"] = radiance_scale * np.exp(rng.normal()) * rng.uniform(0.5, 1.0, 3)
        m["ior"] = 1.0
        new_mats.append(m)
        new_tm.append(np.array([len(mats) + k], np.uint32))
    return (np.concatenate(new_pos).astype(np.float32), np.concatenate(new_idx),
            np.concatenate(new_tm), np.concatenate(new_mats))


def instanced_room(num_instances=6, seed=4):
    """The box room (its walls and light as world triangles) with
    num_instances placements of two small displaced spheres inside it, each
    instance with its own Substitute material.  Returns a dict with meshes,
    mesh_of, world_to_object, object_to_world, instance_materials, positions,
    indices, triangle_materials and materials, as the renderer's instanced
    scene takes them."""
    import numpy as np

    from sprout_amd import instances, scenes

    pos, idx, tmat, mats = scenes.room()
    meshes = [scenes.displaced_icosphere(sub, octaves=3, amplitude=0.1, noise_seed=seed + k,
                                         num_random=0) for k, sub in enumerate((2, 3))]
    _, o2w = instances.random_transforms(num_instances, seed, extent=0.55, scale=(0.15, 0.3))
    o2w = o2w.reshape(-1, 4, 4).astype(np.float64)
    o2w[:, 3, 1] += 0.8  # inside the room, y in [0.25, 1.35]
    w2o = np.linalg.inv(o2w)
    rng = np.random.default_rng(seed + 1)
    im = np.zeros(num_instances, scenes.MATERIAL_DTYPE)
    im["kind"] = scenes.SUBSTITUTE
    im["color"] = rng.uniform(0.1, 0.9, (num_instances, 3))
    im["roughness"] = rng.uniform(0.1, 0.9, num_instances)
    im["metallic"] = (rng.random(num_instances) < 0.3).astype(np.float32)
    im["ior"] = 1.46
    return {
        "meshes": meshes,
        "mesh_of": (np.arange(num_instances) % 2).astype(np.uint32),
        "world_to_object": w2o.astype(np.float32).reshape(-1, 16),
        "object_to_world": o2w.astype(np.float32).reshape(-1, 16),
        "instance_materials": (len(mats) + np.arange(num_instances)).astype(np.uint32),
        "positions": pos,
        "indices": idx,
        "triangle_materials": tmat,
        "materials": np.concatenate([mats, im]),
    }


def baked(scene):
    """An instanced scene as one triangle soup: each instance's mesh moved to
    world space with the oracle's transform_point, then the world triangles.
    Returns positions, indices, per-triangle material and materials."""
    import numpy as np

    from oracle import oracle as O

    pos, idx, tmat = [], [], []
    base = 0
    for i, m in enumerate(scene["mesh_of"]):
        p, t = scene["meshes"][m]
        wp = np.array([O.transform_point(scene["object_to_world"][i], v) for v in p], np.float32)
        pos.append(wp)
        idx.append(t + base)
        tmat.append(np.full(len(t), scene["instance_materials"][i], np.uint32))
        base += len(wp)
    pos.append(scene["positions"])
    idx.append(scene["indices"] + base)
    tmat.append(scene["triangle_materials"])
    return (np.concatenate(pos).astype(np.float32), np.concatenate(idx).astype(np.uint32),
            np.concatenate(tmat).astype(np.uint32), scene["materials"])


def oracle_instances(scene):
    """The oracle's form of an instanced scene for O.render(instances=...)."""
    from oracle import oracle as O

    return O.PtInstances(scene["meshes"], scene["mesh_of"], scene["world_to_object"],
                         scene["object_to_world"], scene["instance_materials"], scene["positions"],
                         scene["indices"])
