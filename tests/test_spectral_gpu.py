"""Hero-wavelength transport on the GPU (MI_RENDER_SPECTRAL) against the
oracle's spectral path tracer on the same seeds, and against the GPU's own
rgb renders where the mathematics fixes the answer:

- the box room and the room with a chromatic medium: the same images as the
  oracle's spectral mode, up to the device's last-bit differences;
- the fBm cloud (grey medium, grey ground): the spectral image equals the
  rgb one up to the film's wavelength noise (the paths draw the same random
  numbers, and grey transport does not depend on wavelength);
- graph and kernel-by-kernel launches, and regenerating waves, give the
  same spectral image bit for bit;
- guided spectral renders (medium vertices guided too) converge to the
  plain spectral image.
"""

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


def rel_mse(x, r, eps=1e-2):
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def random_grid(shape, seed, scale=3.0):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    g = gaussian_filter(rng.random(shape), 2.0)
    g = (g - g.min()) / (g.max() - g.min())
    g[g < 0.3] = 0.0
    return (g * scale).astype(np.float32)


def image(gpu, sc, cam, spp, **kw):
    import torch

    acc, _ = gpu.render(sc, cam, spp, **kw)
    torch.cuda.synchronize()
    return acc.cpu().numpy() / spp


def test_room_same_seeds_as_oracle(gpu):
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    w, h, spp = 64, 48, 16
    log(f"spectral box room {w}x{h} {spp} spp against the oracle")
    cam = gpu.camera(w, h, **scenes.ROOM_CAMERA)
    img = image(gpu, sc, cam, spp, samples_per_wave=4, seed=5, spectral=True)
    ref = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA,
                   light_tree(pos, idx, tmat, mats), seed=5, spectral=True) / spp
    assert np.all(np.isfinite(img))
    diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
    log(f"  pixels within 1e-4: {np.mean(diff < 1e-4):.4f}, relMSE {rel_mse(img, ref):.3g}")
    assert np.mean(diff < 1e-4) > 0.95
    assert rel_mse(img, ref) < 1e-5
    # and it is not the rgb image: the chromatic walls' interreflection differs
    rgb = image(gpu, sc, cam, spp, samples_per_wave=4, seed=5)
    assert not np.allclose(img, rgb, rtol=1e-3)
    sc.release()


def test_room_with_chromatic_medium_same_seeds_as_oracle(gpu):
    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    grid = random_grid((40, 48, 56), 1)
    box = ((-0.6, 0.3, -0.6), (0.6, 1.5, 0.6))
    cc_a, cc_s, g = (0.3, 0.4, 0.5), (1.2, 0.9, 0.6), 0.5
    M.set_medium(sc, grid, *box, cc_a, cc_s, g)
    om = O.Medium(grid, *box, cc_a, cc_s, g)
    w, h, spp = 48, 40, 16
    log("spectral room with a chromatic medium against the oracle")
    cam = gpu.camera(w, h, **scenes.ROOM_CAMERA)
    img = image(gpu, sc, cam, spp, samples_per_wave=4, seed=8, spectral=True)
    ref = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA,
                   light_tree(pos, idx, tmat, mats), seed=8, medium=om, spectral=True) / spp
    assert np.isfinite(img).all()
    diff = np.abs(img - ref) / np.maximum(ref, 1e-3)
    log(f"  pixels within 1e-4: {np.mean(diff < 1e-4):.4f}, relMSE {rel_mse(img, ref):.3g}")
    # as in rgb: tracking turns last-bit differences of log / exp into other
    # decisions for a few paths
    assert np.mean(diff < 1e-4) > 0.9
    assert rel_mse(img, ref) < 1e-3
    sc.release()


@pytest.fixture(scope="module")
def cloud(gpu):
    from sprout_amd import medium as M
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.cloud_scene()
    sc = gpu.Scene(pos, idx, tmat, mats)
    grid = scenes.fbm_density(64, 6, 1, 20.0)
    M.set_medium(sc, grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S,
                 scenes.CLOUD_G, majorants="flat")
    yield sc
    sc.release()


def test_grey_cloud_spectral_equals_rgb(gpu, cloud):
    from sprout_amd import scenes

    w, h, spp = 48, 48, 64
    cam = gpu.camera(w, h, **scenes.CLOUD_CAMERA)
    log("fBm cloud 64^3 (grey): spectral against rgb, same seeds")
    rgb = image(gpu, cloud, cam, spp, samples_per_wave=16, seed=2)
    spec = image(gpu, cloud, cam, spp, samples_per_wave=16, seed=2, spectral=True)
    assert np.isfinite(spec).all()
    k = 4
    blocks = lambda x: x.reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))  # noqa: E731
    e = rel_mse(blocks(spec), blocks(rgb))
    log(f"  relMSE of 4x4 blocks {e:.3g}, means {spec.mean():.5g} {rgb.mean():.5g}")
    assert abs(spec.mean() / rgb.mean() - 1) < 0.01
    assert e < 1e-3


def test_spectral_launch_modes_agree(gpu, cloud):
    import torch

    from sprout_amd import scenes

    cam = gpu.camera(32, 32, **scenes.CLOUD_CAMERA)
    log("fBm cloud 64^3 spectral: graph, launches, regenerating waves")
    kw = dict(samples_per_wave=4, seed=3, spectral=True)
    a, _ = gpu.render(cloud, cam, 8, **kw)
    b, _ = gpu.render(cloud, cam, 8, graph=False, **kw)
    c, _ = gpu.render(cloud, cam, 8, regenerate=True, **kw)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(a, c)


def test_guided_spectral_converges_to_plain(gpu, cloud):
    """the guide at surface and medium vertices, trained online, in spectral
    mode: the same expectation as the plain spectral render"""
    import torch

    from sprout_amd import guide, scenes

    g = guide.Guide(9)
    cam = gpu.camera(32, 32, **scenes.CLOUD_CAMERA)
    log("fBm cloud 64^3 spectral: guided (medium vertices too) vs plain, 2048 spp")
    ref = image(gpu, cloud, cam, 2048, samples_per_wave=256, seed=41, spectral=True)
    acc, st = gpu.render_guided(cloud, cam, 2048, g, fraction=0.5, train=True, train_steps=2,
                                samples_per_wave=256, seed=42, stats=True, guide_medium=True,
                                spectral=True)
    torch.cuda.synchronize()
    img = acc.cpu().numpy() / 2048
    blocks = lambda x: x.reshape(8, 4, 8, 4, 3).mean(axis=(1, 3))  # noqa: E731
    e = rel_mse(blocks(img), blocks(ref))
    log(f"  relMSE of 4x4 blocks {e:.3g}, means {img.mean():.5g} {ref.mean():.5g}")
    assert np.isfinite(img).all()
    assert st.trained > 0
    assert e < 3e-3
    assert abs(img.mean() / ref.mean() - 1) < 0.02
    g.release()
