"""Participating medium on the GPU against the oracle: batch delta tracking
and ratio tracking with the same generators (mi_medium_track vs
oracle_medium_batch), renders of the room and the cloud scene with a medium
with the same random numbers as the oracle's path tracer, argument checks,
and the full-size 256^3 fBm cloud (finite, converging)."""

import numpy as np
import pytest

from oracle import oracle as O
from scene_util import light_tree

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


def random_grid(shape, seed, scale=3.0):
    rng = np.random.default_rng(seed)
    from scipy.ndimage import gaussian_filter

    g = gaussian_filter(rng.random(shape), 2.0)
    g = (g - g.min()) / (g.max() - g.min())
    g[g < 0.3] = 0.0
    return (g * scale).astype(np.float32)


def random_rays(n, seed, box_lo, box_hi):
    """rays from outside and inside the box, some ending inside it"""
    rng = np.random.default_rng(seed)
    lo, hi = np.array(box_lo), np.array(box_hi)
    c = 0.5 * (lo + hi)
    ext = hi - lo
    o = np.where(rng.random((n, 1)) < 0.5, c + (rng.random((n, 3)) - 0.5) * ext * 3.0,
                 lo + rng.random((n, 3)) * ext)
    target = lo + rng.random((n, 3)) * ext
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = np.where(rng.random(n) < 0.3, rng.random(n) * 3.0, 3.4e38)
    return r


@pytest.fixture(scope="module")
def medium_scene(gpu):
    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    grid = random_grid((40, 48, 56), 1)
    box = ((-0.6, 0.3, -0.6), (0.6, 1.5, 0.6))
    cc_a, cc_s, g = (0.3, 0.4, 0.5), (1.2, 0.9, 0.6), 0.5
    M.set_medium(sc, grid, *box, cc_a, cc_s, g)
    om = O.Medium(grid, *box, cc_a, cc_s, g)
    yield sc, om, (pos, idx, tmat, mats), box
    sc.release()


@pytest.mark.parametrize("transmittance", [False, True])
@pytest.mark.parametrize("nearest", [False, True])
@pytest.mark.parametrize("depth", [0, 30])
def test_batch_tracking_matches_oracle(gpu, medium_scene, transmittance, nearest, depth):
    from sprout_amd import medium as M

    sc, om, _, box = medium_scene
    rays = random_rays(20000, 5 + depth, *box)
    log(f"medium batch: transmittance {transmittance}, nearest {nearest}, depth {depth}")
    dev = M.track(sc, rays, depth=depth, seed=9, transmittance=transmittance, nearest=nearest)
    ref = om.batch(rays, depth=depth, seed=9, transmittance=transmittance, nearest=nearest)
    same_event = dev[:, 0] == ref[:, 0]
    assert same_event.mean() > 0.995, same_event.mean()
    close = same_event & np.all(np.abs(dev[:, 1:] - ref[:, 1:]) <= 1e-4 * np.abs(ref[:, 1:]) + 1e-6,
                                axis=1)
    assert close.mean() > 0.99, close.mean()
    # the rays did meet the medium
    if transmittance:
        assert (ref[:, 2] < 1).mean() > 0.2
    else:
        assert (ref[:, 0] == 1).mean() > 0.1


@pytest.mark.parametrize("transmittance", [False, True])
def test_flat_majorants_match_oracle(gpu, transmittance):
    """The flat majorant grid (MI_MAJORANTS_FLAT) on a cloud of two
    contrasts, one under the gridtree's split threshold: batch tracking
    matches the oracle's over the same flat grid, and its transmittance
    agrees in the mean with the gridtree's (both are unbiased)."""
    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    grid = random_grid((64, 64, 64), 2)
    grid[:, :, :32] *= np.float32(0.03)  # a faint half
    box = ((-0.6, 0.3, -0.6), (0.6, 1.5, 0.6))
    cc_a, cc_s, g = (0.3, 0.4, 0.5), (1.2, 0.9, 0.6), 0.5
    flat = gpu.Scene(pos, idx, tmat, mats)
    tree = gpu.Scene(pos, idx, tmat, mats)
    M.set_medium(flat, grid, *box, cc_a, cc_s, g, majorants="flat")
    M.set_medium(tree, grid, *box, cc_a, cc_s, g)
    with pytest.raises(ValueError):
        M.set_medium(tree, grid, *box, cc_a, cc_s, g, majorants="octree")
    om = O.Medium(grid, *box, cc_a, cc_s, g, majorants="flat")
    rays = random_rays(20000, 21, *box)
    log(f"flat majorants batch: transmittance {transmittance}")
    dev = M.track(flat, rays, seed=4, transmittance=transmittance)
    ref = om.batch(rays, seed=4, transmittance=transmittance)
    same_event = dev[:, 0] == ref[:, 0]
    assert same_event.mean() > 0.995, same_event.mean()
    close = same_event & np.all(np.abs(dev[:, 1:] - ref[:, 1:]) <= 1e-4 * np.abs(ref[:, 1:]) + 1e-6,
                                axis=1)
    assert close.mean() > 0.99, close.mean()
    if transmittance:
        n = 200000
        many = np.repeat(rays[:16], n // 16, axis=0)
        a = M.track(flat, many, seed=5, transmittance=True)[:, 2:].reshape(16, -1, 3)
        b = M.track(tree, many, seed=6, transmittance=True)[:, 2:].reshape(16, -1, 3)
        sem = np.sqrt(a.var(axis=1) / a.shape[1] + b.var(axis=1) / b.shape[1])
        assert np.all(np.abs(a.mean(axis=1) - b.mean(axis=1)) <= 5 * sem + 1e-4)
    flat.release()
    tree.release()


def test_room_with_medium_same_seeds_as_oracle(gpu, medium_scene):
    import torch

    from sprout_amd import scenes

    sc, om, scene, _ = medium_scene
    w, h, spp = 48, 40, 16
    log("room with a medium, same seeds as the oracle")
    cam = gpu.camera(w, h, **scenes.ROOM_CAMERA)
    img, stats = gpu.render(sc, cam, spp, samples_per_wave=4, seed=8)
    torch.cuda.synchronize()
    img = img.cpu().numpy() / spp
    ref = O.render(*scene, w, h, spp, scenes.ROOM_CAMERA, light_tree(*scene), seed=8,
                   medium=om) / spp
    assert np.isfinite(img).all()
    diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
    # tracking amplifies last-bit differences of log / exp into different
    # decisions for a few paths; nearly every pixel is still identical
    assert np.mean(diff < 1e-4) > 0.9, np.mean(diff < 1e-4)
    assert float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2))) < 1e-3


def test_cloud_scene_graph_and_launches_agree(gpu):
    import torch

    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.cloud_scene()
    sc = gpu.Scene(pos, idx, tmat, mats)
    try:
        grid = scenes.fbm_density(32, 6, 1, 20.0)
        M.set_medium(sc, grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S,
                     scenes.CLOUD_G)
        om = O.Medium(grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S, scenes.CLOUD_G)
        w, h, spp = 32, 32, 8
        cam = gpu.camera(w, h, **scenes.CLOUD_CAMERA)
        log("cloud 32^3: graph vs launches")
        a, _ = gpu.render(sc, cam, spp, samples_per_wave=4, seed=3)
        b, _ = gpu.render(sc, cam, spp, samples_per_wave=4, seed=3, graph=False)
        c, _ = gpu.render(sc, cam, spp, samples_per_wave=4, seed=3, regenerate=True)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        # regenerating waves: tracking reads each path's own depth
        assert torch.equal(a, c)
        log("cloud 32^3: oracle, same seeds")
        img = a.cpu().numpy() / spp
        ref = O.render(pos, idx, tmat, mats, w, h, spp, scenes.CLOUD_CAMERA,
                       light_tree(pos, idx, tmat, mats), seed=3, medium=om) / spp
        diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
        assert np.mean(diff < 1e-4) > 0.8, np.mean(diff < 1e-4)
        assert float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2))) < 2e-3
    finally:
        sc.release()


def test_full_size_cloud_is_finite_and_converges(gpu):
    """256^3 fBm cloud (max density 20): two independent renders agree to
    within their noise and every pixel is finite"""
    import torch

    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.cloud_scene()
    sc = gpu.Scene(pos, idx, tmat, mats)
    try:
        log("cloud 256^3: building the medium")
        grid = scenes.fbm_density(256, 6, 1, 20.0)
        assert grid.max() > 15.0 and (grid == 0).mean() > 0.3
        M.set_medium(sc, grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S,
                     scenes.CLOUD_G)
        w, h, spp = 64, 64, 64
        cam = gpu.camera(w, h, **scenes.CLOUD_CAMERA)
        log("cloud 256^3: two renders")
        a, _ = gpu.render(sc, cam, spp, samples_per_wave=16, seed=1)
        b, _ = gpu.render(sc, cam, spp, samples_per_wave=16, seed=2)
        torch.cuda.synchronize()
        a = a.cpu().numpy() / spp
        b = b.cpu().numpy() / spp
        assert np.isfinite(a).all() and np.isfinite(b).all()
        assert a.mean() > 0
        assert abs(a.mean() - b.mean()) < 0.02 * a.mean()
    finally:
        sc.release()


def test_bad_medium_arguments(gpu):
    from sprout_amd import _lib
    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    try:
        grid = np.ones((4, 4, 4), np.float32)
        with pytest.raises(_lib.MiError):
            M.set_medium(sc, grid, (0, 0, 0), (0, 1, 1), (1, 1, 1), (1, 1, 1), 0.0)  # empty box
        with pytest.raises(_lib.MiError):
            M.set_medium(sc, grid, (0, 0, 0), (1, 1, 1), (1, 1, 1), (1, 1, 1), 1.0)  # g = 1
        with pytest.raises(_lib.MiError):
            M.set_medium(sc, -grid, (0, 0, 0), (1, 1, 1), (1, 1, 1), (1, 1, 1), 0.0)
        with pytest.raises(_lib.MiError):
            M.set_medium(sc, grid * 1e5, (0, 0, 0), (1, 1, 1), (1, 1, 1), (1, 1, 1), 0.0)
        with pytest.raises(_lib.MiError):
            M.track(sc, np.zeros((4, 8), np.float32))  # no medium
        M.set_medium(sc, grid, (0, 0, 0), (1, 1, 1), (1, 1, 1), (1, 1, 1), 0.0)
        M.clear_medium(sc)
        with pytest.raises(_lib.MiError):
            M.track(sc, np.zeros((4, 8), np.float32))
    finally:
        sc.release()
