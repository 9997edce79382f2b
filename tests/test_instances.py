"""Instanced meshes: the oracle's two-level restatement (oracle/instances.c)
pinned before it checks the GPU.

- transform_point / transform_vector / AABB::transform equal the reference's
  own functions compiled here (oracle/_ref), bit for bit;
- the top-level tree the oracle builds over degenerate triangles (lo, hi, lo)
  equals the tree the reference's scene::bvh::Builder builds over the same
  boxes, node for node and slot for slot;
- the two-level closest hit equals brute force over every instance and
  triangle in object space (the same primitive test), and any-hit agrees with
  it.
"""

import numpy as np
import pytest

from oracle import oracle as O


def meshes(seed=1):
    from sprout_amd import scenes

    out = []
    for k, sub in enumerate((2, 3)):
        pos, idx = scenes.displaced_icosphere(sub, octaves=4, amplitude=0.15, noise_seed=seed + k,
                                              num_random=0)
        out.append((pos, idx, O.build(pos, idx)))
    return out


def transforms(n, seed):
    from sprout_amd import instances

    return instances.random_transforms(n, seed, extent=3.0)


def rays_at(n, seed, extent=4.0):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 9.0
    target = rng.uniform(-extent, extent, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = np.where(rng.random(n) < 0.2, rng.uniform(5, 12, n), 3.4e38)
    return r


@pytest.fixture(scope="module")
def ref():
    l = O.ref_prop_lib()
    if l is None:
        pytest.skip("the reference library is not built here (oracle/ref.mk)")
    return l


def test_transforms_equal_the_reference(ref):
    import ctypes

    rng = np.random.default_rng(3)
    w2o, o2w = transforms(200, 4)
    for i in range(200):
        m = np.ascontiguousarray(w2o[i], np.float32)
        v = rng.normal(size=3).astype(np.float32) * 10
        a = np.zeros(3, np.float32)
        ref.sref_transform_point(m.ctypes.data, v.ctypes.data, a.ctypes.data)
        assert np.array_equal(a.view(np.uint32), O.transform_point(m, v).view(np.uint32))
        ref.sref_transform_vector(m.ctypes.data, v.ctypes.data, a.ctypes.data)
        assert np.array_equal(a.view(np.uint32), O.transform_vector(m, v).view(np.uint32))
        lo = rng.normal(size=3).astype(np.float32)
        hi = lo + rng.random(3).astype(np.float32)
        b = np.zeros(6, np.float32)
        ref.sref_transform_box(np.ascontiguousarray(o2w[i]).ctypes.data, lo.ctypes.data,
                               hi.ctypes.data, b.ctypes.data)
        assert np.array_equal(b.view(np.uint32), O.transform_box(lo, hi, o2w[i]).view(np.uint32))
    del ctypes


@pytest.mark.parametrize("n,seed", [(1, 1), (7, 2), (64, 3), (300, 4)])
def test_top_level_tree_equals_the_reference_prop_builder(ref, n, seed):
    ms = meshes()
    w2o, o2w = transforms(n, seed)
    mesh_of = np.arange(n, dtype=np.uint32) % len(ms)
    inst = O.Instances(ms, mesh_of, w2o, o2w)
    # the instance boxes: each mesh root box transformed
    for i in range(n):
        root = ms[mesh_of[i]][2].nodes[0]
        assert np.array_equal(inst.boxes[i], O.transform_box(root["lo"], root["hi"], o2w[i]))
    nodes, idx = O.ref_build_props(ref, inst.boxes)
    assert O.nodes_equal(nodes, inst.top_nodes)
    assert np.array_equal(idx, inst.top_slots)


def brute(ms, mesh_of, w2o, rays):
    """closest over all instances and triangles in object space (min t, later
    instance wins a tie): (t, triangle, instance) per ray"""
    n = len(rays)
    best_t = np.array(rays[:, 7], np.float32)
    tri = np.full(n, 0xFFFFFFFF, np.uint32)
    ins = np.full(n, 0xFFFFFFFF, np.uint32)
    for p in range(len(mesh_of)):
        pos, idx, _ = ms[mesh_of[p]]
        obj = rays.copy()
        for r in range(n):
            obj[r, 0:3] = O.transform_point(w2o[p], rays[r, 0:3])
            obj[r, 4:7] = O.transform_vector(w2o[p], rays[r, 4:7])
        obj[:, 7] = best_t
        h = O.brute_force(pos, idx, obj)
        t = h[:, 0].view(np.float32)
        hit = h[:, 3] != 0xFFFFFFFF
        best_t = np.where(hit, t, best_t)
        tri = np.where(hit, h[:, 3], tri)
        ins = np.where(hit, p, ins)
    return best_t, tri, ins


def test_two_level_closest_hit_equals_brute_force():
    ms = meshes(5)
    n = 12
    w2o, o2w = transforms(n, 6)
    mesh_of = (np.arange(n) % 2).astype(np.uint32)
    inst = O.Instances(ms, mesh_of, w2o, o2w)
    rays = rays_at(600, 7)
    hits = inst.trace(rays)
    t_b, tri_b, ins_b = brute(ms, mesh_of, w2o, rays)
    hit = hits[:, 4] != 0xFFFFFFFF
    assert hit.mean() > 0.3
    assert np.array_equal(hit, ins_b != 0xFFFFFFFF)
    assert np.array_equal(hits[:, 0].view(np.float32)[hit], t_b[hit])
    # instance and triangle agree except on exact ties of t
    same = (hits[:, 3] == tri_b) & (hits[:, 4] == ins_b)
    assert same[hit].mean() > 0.995
    occ = inst.trace(rays, any_hit=True)
    assert np.array_equal(occ.astype(bool), hit)
