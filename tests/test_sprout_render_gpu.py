"""The GPU path tracer with shading normals (mi_scene_set_normals) against
the reference's own renderer and against the oracle.

* In expectation: the takes of tests/golden/take_fixture, take_smooth and
  take_room (the benchmark's box room) rendered at 16 384 spp against the
  reference's 65 536-spp renders (tests/golden/sprout_renders.npz, made by
  tests/golden/make_sprout_renders.py): image means within 0.5 %, 16 x 16
  blocks within 1 % (small, room) and 2.5 % (smooth; its residual sits on
  the sphere's silhouette, see tests/test_sprout_render.py).
* Under the same random numbers: the smooth take against the oracle.
* The entry point: normals of the wrong count or without direction are
  refused, and None goes back to the geometric normal.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O
from scene_util import light_tree

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
TAKES = {"small": os.path.join(GOLDEN, "take_fixture", "small.take"),
         "smooth": os.path.join(GOLDEN, "take_smooth", "smooth.take"),
         "room": os.path.join(GOLDEN, "take_room", "room.take")}
TOLERANCE = {"small": (0.005, 0.01), "smooth": (0.005, 0.025), "room": (0.005, 0.01)}


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh, render

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield render
    torch.cuda.synchronize()


def blocks(img, b):
    h, w, _ = img.shape
    return img[:h // b * b, :w // b * b].reshape(h // b, b, w // b, b, 3).mean(axis=(1, 3))


def gpu_image(render, t, spp, seed, normals=True):
    import torch

    pos, idx, tmat, mats = t.scene
    sc = render.Scene(pos, idx, tmat, mats, normals=t.normals if normals else None)
    cam = render.perspective(t.width, t.height, **t.camera)
    acc, _ = render.render(sc, cam, spp, samples_per_wave=min(spp, 256), seed=seed,
                           min_bounces=t.min_bounces, max_bounces=t.max_bounces)
    torch.cuda.synchronize()
    sc.release()
    return acc.cpu().numpy() / spp


@pytest.mark.parametrize("name", sorted(TAKES))
def test_gpu_converges_to_sprouts_render(gpu, name):
    from sprout_amd import take

    with np.load(os.path.join(GOLDEN, "sprout_renders.npz"), allow_pickle=False) as z:
        ref = z[name]
    t = take.Take(TAKES[name])
    print(f"take {name}: {len(t.scene[1])} triangles, 16384 spp on the GPU", flush=True)
    img = gpu_image(gpu, t, 16384, seed=19)
    mean = np.abs(img.mean(axis=(0, 1)) / ref.mean(axis=(0, 1)) - 1).max()
    rb, ib = blocks(ref, 16), blocks(img, 16)
    block = (np.abs(ib - rb) / np.maximum(rb, 2e-3)).max()
    tol_mean, tol_block = TOLERANCE[name]
    assert mean < tol_mean, mean
    assert block < tol_block, block
    if name != "room":  # the room's flat walls need no shading normals
        # the geometric normal instead: far off
        flat = gpu_image(gpu, t, 4096, seed=19, normals=False)
        fb = blocks(flat, 16)
        assert (np.abs(fb - rb) / np.maximum(rb, 2e-3)).max() > 4 * tol_block


def test_filtered_sensor_converges_to_sprouts_render(gpu):
    """the fixture take through its Gaussian sensor at 16 384 spp against the
    reference's Filtered_2p0 sensor (see tests/test_sprout_render.py for the
    top-edge rows beside the lamp)"""
    import torch

    from sprout_amd import take

    with np.load(os.path.join(GOLDEN, "sprout_renders.npz"), allow_pickle=False) as z:
        ref = z["small_filtered"]
    t = take.Take(TAKES["small"])
    pos, idx, tmat, mats = t.scene
    sc = gpu.Scene(pos, idx, tmat, mats, normals=t.normals)
    cam = gpu.perspective(t.width, t.height, **t.camera)
    acc, _ = gpu.render_filtered(sc, cam, 16384, gpu.sensor(), samples_per_wave=256, seed=23,
                                 min_bounces=t.min_bounces, max_bounces=t.max_bounces)
    torch.cuda.synchronize()
    sc.release()
    acc = acc.cpu().numpy()
    img = acc[..., :3] / acc[..., 3:4]
    mean = np.abs(img.mean(axis=(0, 1)) / ref.mean(axis=(0, 1)) - 1).max()
    rb, ib = blocks(ref, 16), blocks(img, 16)
    block = (np.abs(ib - rb) / np.maximum(rb, 2e-3)).max()
    assert mean < 0.006, mean
    assert block < 0.015, block
    inner = img[4:-4, 4:-4].mean() / ref[4:-4, 4:-4].mean() - 1
    assert abs(inner) < 0.002, inner


def test_smooth_take_equals_the_oracle(gpu):
    """same seeds: the interpolated normals and the frames built from them
    agree with the oracle's up to the device's float rounding"""
    from sprout_amd import take

    t = take.Take(TAKES["smooth"])
    pos, idx, tmat, mats = t.scene
    spp = 64
    img = gpu_image(gpu, t, spp, seed=3)
    ref = O.render(pos, idx, tmat, mats, t.width, t.height, spp, t.camera,
                   light_tree(pos, idx, tmat, mats), seed=3, min_bounces=t.min_bounces,
                   max_bounces=t.max_bounces, normals=t.normals) / spp
    assert img.mean() > 1e-3
    diff = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.mean(diff < 1e-4) > 0.95, np.mean(diff < 1e-4)
    assert np.mean((img - ref) ** 2 / (ref * ref + 1e-2)) < 1e-5


def test_normals_in_regenerating_and_spectral_waves(gpu):
    """the smooth take with shading normals: regenerating waves give the
    wave-by-wave image bit for bit, and hero-wavelength transport matches
    the oracle's spectral mode under the same seeds"""
    import torch

    from sprout_amd import take

    t = take.Take(TAKES["smooth"])
    pos, idx, tmat, mats = t.scene
    sc = gpu.Scene(pos, idx, tmat, mats, normals=t.normals)
    cam = gpu.perspective(t.width, t.height, **t.camera)
    kw = dict(samples_per_wave=4, seed=3, min_bounces=t.min_bounces, max_bounces=t.max_bounces)
    a, _ = gpu.render(sc, cam, 16, **kw)
    b, _ = gpu.render(sc, cam, 16, regenerate=True, **kw)
    c, _ = gpu.render(sc, cam, 16, spectral=True, **kw)
    torch.cuda.synchronize()
    sc.release()
    a, b, c = (x.cpu().numpy() / 16 for x in (a, b, c))
    assert a.mean() > 1e-3 and np.array_equal(a, b)
    ref = O.render(pos, idx, tmat, mats, t.width, t.height, 16, t.camera,
                   light_tree(pos, idx, tmat, mats), seed=3, min_bounces=t.min_bounces,
                   max_bounces=t.max_bounces, normals=t.normals, spectral=True) / 16
    diff = np.abs(c - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.mean(diff < 1e-4) > 0.95, np.mean(diff < 1e-4)
    assert np.mean((c - ref) ** 2 / (ref * ref + 1e-2)) < 1e-5


def test_guided_render_with_normals_converges_to_sprouts_render(gpu):
    """the guided integrator on the smooth take (the guide trained online at
    the first two vertices, mixture fraction 0.5) stays unbiased with shading
    normals: it converges to the reference's render like the plain one"""
    import torch

    from sprout_amd import guide, take

    with np.load(os.path.join(GOLDEN, "sprout_renders.npz"), allow_pickle=False) as z:
        ref = z["smooth"]
    t = take.Take(TAKES["smooth"])
    pos, idx, tmat, mats = t.scene
    sc = gpu.Scene(pos, idx, tmat, mats, normals=t.normals)
    cam = gpu.perspective(t.width, t.height, **t.camera)
    g = guide.Guide(8)
    spp = 8192
    acc, _ = gpu.render_guided(sc, cam, spp, g, fraction=0.5, train=True, train_steps=2,
                               batch=16384, learning_rate=3e-3, record_depth=2, guide_depth=2,
                               samples_per_wave=256, seed=29, min_bounces=t.min_bounces,
                               max_bounces=t.max_bounces)
    torch.cuda.synchronize()
    g.release()
    sc.release()
    img = acc.cpu().numpy() / spp
    mean = np.abs(img.mean(axis=(0, 1)) / ref.mean(axis=(0, 1)) - 1).max()
    rb, ib = blocks(ref, 16), blocks(img, 16)
    block = (np.abs(ib - rb) / np.maximum(rb, 2e-3)).max()
    assert mean < 0.01, mean
    assert block < 0.04, block


def test_set_normals_arguments(gpu):
    from sprout_amd import _lib, take

    t = take.Take(TAKES["small"])
    pos, idx, tmat, mats = t.scene
    sc = gpu.Scene(pos, idx, tmat, mats)
    with pytest.raises(ValueError):
        sc.set_normals(t.normals[:-1])
    bad = t.normals.copy()
    bad[1] = 0
    with pytest.raises(_lib.MiError):
        sc.set_normals(bad)
    sc.set_normals(t.normals)
    sc.set_normals(None)
    sc.release()
