"""The wavefront path tracer over mesh instances (mi_scene_create_instanced,
mi_render) against the oracle's instanced path tracer on the same seeds, and
against the GPU's own render of the same scene baked into one soup.

Both the GPU and the oracle draw the same random numbers per path and form
hit points and normals with the same operations, so with equal seeds the
images agree up to the device's sin/cos ulps, as for triangle scenes.
"""

import numpy as np
import pytest

from oracle import oracle as O
from scene_util import baked, instanced_room, light_tree, oracle_instances

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh, render

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield render
    torch.cuda.synchronize()


def log(msg):
    print(msg, flush=True)


def rel_mse(x, r, eps=1e-2):
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def make(gpu, sc):
    return gpu.InstancedScene(sc["meshes"], sc["mesh_of"], sc["world_to_object"],
                              sc["object_to_world"], sc["instance_materials"], sc["positions"],
                              sc["indices"], sc["triangle_materials"], sc["materials"])


@pytest.fixture(scope="module")
def room(gpu):
    sc = instanced_room()
    s = make(gpu, sc)
    yield s, sc
    s.release()


def image(gpu, s, w, h, spp, **kw):
    import torch

    from sprout_amd import scenes

    cam = gpu.camera(w, h, **scenes.ROOM_CAMERA)
    acc, _ = gpu.render(s, cam, spp, **kw)
    torch.cuda.synchronize()
    return acc.cpu().numpy() / spp


@pytest.mark.parametrize("spw,graph", [(4, True), (2, False)])
def test_same_seeds_as_oracle(gpu, room, spw, graph):
    from sprout_amd import scenes

    s, sc = room
    w, h, spp = 64, 48, 16
    log(f"instanced room {w}x{h} {spp} spp, waves of {spw}, graph {graph}")
    img = image(gpu, s, w, h, spp, samples_per_wave=spw, graph=graph, seed=5)
    tree = light_tree(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"])
    ref = O.render(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"], w, h,
                   spp, scenes.ROOM_CAMERA, tree, seed=5, instances=oracle_instances(sc)) / spp
    assert np.all(np.isfinite(img))
    diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
    assert np.mean(diff < 1e-4) > 0.95, np.mean(diff < 1e-4)
    assert rel_mse(img, ref) < 1e-5


def test_equals_the_baked_soup(gpu, room):
    s, sc = room
    log("instanced room against its baked soup, 64x48 32 spp")
    pos, idx, tm, mats = baked(sc)
    soup = gpu.Scene(pos, idx, tm, mats)
    a = image(gpu, s, 64, 48, 32, samples_per_wave=8, seed=9)
    b = image(gpu, soup, 64, 48, 32, samples_per_wave=8, seed=9)
    soup.release()
    diff = np.abs(a - b) / np.maximum(b, 1e-3)
    assert np.mean(diff < 1e-3) > 0.95
    assert rel_mse(a, b) < 1e-4


def test_rejects_bad_scenes(gpu):
    from sprout_amd import _lib

    sc = instanced_room(2)
    bad = dict(sc)
    bad["instance_materials"] = np.array([3, 3], np.uint32)  # the room's light
    with pytest.raises(_lib.MiError):
        make(gpu, bad)
    bad = dict(sc)
    bad["mesh_of"] = np.array([0, 2], np.uint32)
    with pytest.raises(_lib.MiError):
        make(gpu, bad)
    bad = dict(sc)
    bad["instance_materials"] = np.array([0, 99], np.uint32)
    with pytest.raises(_lib.MiError):
        make(gpu, bad)


def test_update_instances_equals_a_fresh_scene(gpu):
    """mi_scene_update_instances rebuilds the top level on the device: a
    render after moving the instances equals, bit for bit, the render of a
    scene created with the new matrices; moving them back gives the first
    image again, and a rejected update keeps the placement."""
    from sprout_amd import _lib

    sc = instanced_room(6, seed=4)
    moved = instanced_room(6, seed=11)  # other placements (its meshes unused)
    s = make(gpu, sc)
    kw = dict(samples_per_wave=4, seed=13)
    log("instanced room 64x48 16 spp, then the instances moved")
    first = image(gpu, s, 64, 48, 16, **kw)
    s.update_instances(moved["world_to_object"], moved["object_to_world"])
    after = image(gpu, s, 64, 48, 16, **kw)
    fresh_sc = dict(sc, world_to_object=moved["world_to_object"],
                    object_to_world=moved["object_to_world"])
    fresh = make(gpu, fresh_sc)
    want = image(gpu, fresh, 64, 48, 16, **kw)
    fresh.release()
    assert not np.array_equal(after, first)
    assert np.array_equal(after, want)

    s.update_instances(sc["world_to_object"], sc["object_to_world"])
    assert np.array_equal(image(gpu, s, 64, 48, 16, **kw), first)

    with pytest.raises(_lib.MiError):
        s.update_instances(sc["world_to_object"][:5], sc["object_to_world"][:5])
    bad = sc["object_to_world"].copy()
    bad[2, 0] = np.nan
    with pytest.raises(_lib.MiError):
        s.update_instances(sc["world_to_object"], bad)
    assert np.array_equal(image(gpu, s, 64, 48, 16, **kw), first)
    s.release()

    pos, idx, tm, mats = baked(sc)
    soup = gpu.Scene(pos, idx, tm, mats)
    with pytest.raises(_lib.MiError):
        _lib.check(_lib.load().mi_scene_update_instances(
            soup.handle, sc["world_to_object"].ctypes.data, sc["object_to_world"].ctypes.data, 6),
            "mi_scene_update_instances")
    soup.release()


@pytest.fixture(scope="module")
def batch(gpu):
    """The batched path-tracing scene at full size: 64 placements of four
    20 480-triangle displaced spheres with per-instance Substitute materials
    (roughness U[0.05, 1], metallic Bernoulli(0.3), colour U[0.1, 0.9]^3),
    a ground quad and 10 000 emissive triangles of log-normal power."""
    from sprout_amd import scenes

    sc = scenes.batch_scene_instanced()
    s = make(gpu, sc)
    yield s, sc
    s.release()


def batch_image(gpu, s, w, h, spp, **kw):
    import torch

    from sprout_amd import scenes

    cam = gpu.camera(w, h, **scenes.BATCH_CAMERA)
    acc, _ = gpu.render(s, cam, spp, min_bounces=3, max_bounces=16, **kw)
    torch.cuda.synchronize()
    return acc.cpu().numpy() / spp


def test_batch_scene_same_seeds_as_oracle(gpu, batch):
    """The batched workload at reduced resolution (240 x 135, 16 spp, Russian
    roulette from depth 3, max depth 16, adaptive light-tree splitting over
    the 10 000 emitters) against the oracle's instanced path tracer on the
    same seeds."""
    from sprout_amd import scenes

    s, sc = batch
    w, h, spp = 240, 135, 16
    log(f"batch scene {w}x{h} {spp} spp against the oracle")
    img = batch_image(gpu, s, w, h, spp, samples_per_wave=4, seed=5)
    tree = light_tree(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"])
    ref = O.render(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"], w, h,
                   spp, scenes.BATCH_CAMERA, tree, seed=5, min_bounces=3, max_bounces=16,
                   instances=oracle_instances(sc)) / spp
    assert np.all(np.isfinite(img))
    diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
    close = float(np.mean(diff < 1e-4))
    e = rel_mse(img, ref)
    log(f"  pixels within 1e-4: {close:.4f}, relMSE {e:.3g}, means {img.mean():.5g} {ref.mean():.5g}")
    assert close > 0.95
    assert e < 1e-5
    # regenerating waves over the instanced scene: the same image bit for bit
    regen = batch_image(gpu, s, w, h, spp, samples_per_wave=4, seed=5, regenerate=True)
    assert np.array_equal(img, regen)


def test_batch_scene_equals_the_baked_soup(gpu, batch):
    """The same scene baked into one world-space soup (1.32 M triangles under
    one tree) renders the same image on the same seeds up to the rounding of
    the baked vertices."""
    from sprout_amd import scenes

    s, _ = batch
    pos, idx, tm, mats = scenes.batch_scene()
    soup = gpu.Scene(pos, idx, tm, mats)
    log("batch scene against its baked soup, 240x135 16 spp")
    a = batch_image(gpu, s, 240, 135, 16, samples_per_wave=4, seed=9)
    b = batch_image(gpu, soup, 240, 135, 16, samples_per_wave=4, seed=9)
    soup.release()
    diff = np.abs(a - b) / np.maximum(b, 1e-3)
    close = float(np.mean(diff < 1e-3))
    e = rel_mse(a, b)
    log(f"  pixels within 1e-3: {close:.4f}, relMSE {e:.3g}")
    assert close > 0.9
    assert e < 1e-3


def smooth_normals(p, i):
    """area-weighted vertex normals of a mesh (shading normals for the test)"""
    p = np.asarray(p, np.float64).reshape(-1, 3)
    i = np.asarray(i).reshape(-1, 3)
    f = np.cross(p[i[:, 1]] - p[i[:, 0]], p[i[:, 2]] - p[i[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, i[:, k], f)
    return (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)


def test_mesh_normals_same_seeds_as_oracle(gpu):
    """mi_scene_set_mesh_normals: each mesh's vertex normals interpolated in
    object space and moved to world space; the same image as the oracle's
    instanced path tracer with the same normals; None goes back to the
    geometric normal, bit for bit"""
    from sprout_amd import scenes

    sc = instanced_room()
    s = make(gpu, sc)
    w, h, spp = 64, 48, 16
    plain = image(gpu, s, w, h, spp, samples_per_wave=4, seed=5)
    normals = [smooth_normals(p, i) for p, i in sc["meshes"]]
    for k, n in enumerate(normals):
        s.set_mesh_normals(k, n)
    log(f"instanced room with mesh shading normals {w}x{h} {spp} spp")
    img = image(gpu, s, w, h, spp, samples_per_wave=4, seed=5)
    tree = light_tree(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"])
    inst = O.PtInstances(sc["meshes"], sc["mesh_of"], sc["world_to_object"], sc["object_to_world"],
                         sc["instance_materials"], sc["positions"], sc["indices"],
                         mesh_normals=normals)
    ref = O.render(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"], w, h,
                   spp, scenes.ROOM_CAMERA, tree, seed=5, instances=inst) / spp
    assert not np.array_equal(img, plain)
    diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
    assert np.mean(diff < 1e-4) > 0.95, np.mean(diff < 1e-4)
    assert rel_mse(img, ref) < 1e-5
    with pytest.raises(ValueError):
        s.set_mesh_normals(0, normals[0][:-1])
    for k in range(len(normals)):
        s.set_mesh_normals(k, None)
    assert np.array_equal(image(gpu, s, w, h, spp, samples_per_wave=4, seed=5), plain)
    s.release()
