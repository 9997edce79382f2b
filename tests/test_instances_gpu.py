"""Instanced meshes on the GPU (mi_instances_create, mi_trace_instances_*)
against the oracle's two-level restatement, bit for bit: the top-level tree
and instance boxes, and every result word of closest and any hits, on a
small scene and on the micro-benchmark shape of 64 instances of 20 480-triangle
displaced spheres."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield bvh
    torch.cuda.synchronize()


def log(msg):
    print(msg, flush=True)


def rays_at(n, seed, extent, radius):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * radius
    target = rng.uniform(-extent, extent, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = np.where(rng.random(n) < 0.2, rng.uniform(0.5, 2.0, n) * radius, 3.4e38)
    return r


def scene(gpu, subdivisions, num_meshes, n, seed, extent):
    import torch

    from sprout_amd import instances, scenes

    meshes = []
    trees = []
    for k in range(num_meshes):
        pos, idx = scenes.displaced_icosphere(subdivisions[k % len(subdivisions)], octaves=4,
                                              amplitude=0.15, noise_seed=seed + k, num_random=0)
        meshes.append((pos, idx, O.build(pos, idx)))
        trees.append(gpu.Builder_SAH(16, 64, 4).build(
            torch.from_numpy(pos).cuda(), torch.from_numpy(idx.view(np.int32)).cuda()))
    w2o, o2w = instances.random_transforms(n, seed + 100, extent=extent)
    mesh_of = (np.arange(n) % num_meshes).astype(np.uint32)
    inst = instances.Instances([trees[m] for m in mesh_of], w2o, o2w)
    ref = O.Instances(meshes, mesh_of, w2o, o2w)
    return inst, ref, trees


def check_hits(inst, ref, rays):
    import torch

    d_rays = torch.from_numpy(rays).cuda()
    hits = inst.intersect(d_rays).cpu().numpy().view(np.uint32)
    occ = inst.intersect_p(d_rays).cpu().numpy().view(np.uint32)
    want = ref.trace(rays)
    want_any = ref.trace(rays, any_hit=True)
    assert np.array_equal(hits[:, :5], want), np.argwhere(np.any(hits[:, :5] != want, axis=1))[:5]
    assert np.array_equal(occ, want_any)
    return (want[:, 4] != 0xFFFFFFFF).mean()


def test_small_scene_bit_exact(gpu):
    inst, ref, _ = scene(gpu, (2, 3), 2, 12, 1, 3.0)
    nodes, slots, boxes = inst.download()
    assert np.array_equal(boxes.view(np.uint32), ref.boxes.view(np.uint32))
    assert O.nodes_equal(nodes, ref.top_nodes)
    assert np.array_equal(slots, ref.top_slots)
    log("small instanced scene: 20000 rays")
    frac = check_hits(inst, ref, rays_at(20000, 2, 4.0, 9.0))
    assert frac > 0.2
    inst.release()


def test_benchmark_shape_bit_exact(gpu):
    """64 instances of 4 distinct 20 480-triangle spheres (1.3 M instanced
    triangles) spread over a 20-unit cube"""
    inst, ref, _ = scene(gpu, (5,), 4, 64, 7, 10.0)
    log("64 instances: 100000 rays")
    frac = check_hits(inst, ref, rays_at(100000, 8, 10.0, 25.0))
    assert frac > 0.1
    inst.release()


def test_mesh_in_use_cannot_be_released(gpu):
    from sprout_amd import _lib

    inst, ref, trees = scene(gpu, (2,), 1, 3, 3, 2.0)
    with pytest.raises(_lib.MiError):
        _lib.check(_lib.load().mi_bvh_release(trees[0].handle), "mi_bvh_release")
    handle = inst.handle
    inst.release()
    # the released handle is refused, and the mesh can go now
    assert _lib.load().mi_trace_instances_any(handle, None, 0, None, None) == -4  # MI_ERR_HANDLE
    _lib.check(_lib.load().mi_bvh_release(trees[0].handle), "mi_bvh_release")
    trees[0].handle = None
