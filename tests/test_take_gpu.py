"""A take in sprout's formats (tests/golden/take_fixture) rendered on the GPU
through its filtered sensor, against the oracle's samples of the same scene
splatted by the oracle's filter; same seeds, so the images agree up to the
device's sin/cos ulps."""

import os

import numpy as np
import pytest

from oracle import oracle as O
from scene_util import light_tree

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("golden_ratio", [False, True])
def test_take_renders_as_the_oracle(golden_ratio):
    """The take's Perspective camera, colours mapped to ACEScg, sprout's
    vertex normals (the last geometric normal written to each vertex of a
    JSON mesh without normals), the Gaussian sensor; pixel samples from the
    random or the Golden_ratio sampler."""
    import torch

    from sprout_amd import bvh, render, take

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    t = take.Take(os.path.join(HERE, "golden", "take_fixture", "small.take"))
    pos, idx, tmat, mats = t.scene
    print(f"take fixture: {len(idx)} triangles, {t.width}x{t.height}, {t.samples_per_pixel} spp, "
          f"golden ratio {golden_ratio}", flush=True)
    sc = render.Scene(pos, idx, tmat, mats, normals=t.normals)
    cam = render.perspective(t.width, t.height, **t.camera)
    spp = t.samples_per_pixel
    acc, _ = render.render_filtered(sc, cam, spp, render.sensor(), samples_per_wave=8, seed=3,
                                    min_bounces=t.min_bounces, max_bounces=t.max_bounces,
                                    golden_ratio=golden_ratio)
    torch.cuda.synchronize()
    acc = acc.cpu().numpy()
    sc.release()
    f = O.Filter(O.FILTER_GAUSSIAN, 2.0)
    s = O.render(pos, idx, tmat, mats, t.width, t.height, spp, t.camera,
                 light_tree(pos, idx, tmat, mats), seed=3, min_bounces=t.min_bounces,
                 max_bounces=t.max_bounces, border=2, samples=True, golden_ratio=golden_ratio,
                 normals=t.normals)
    ref = O.filtered_sums(s, t.width, t.height, f)
    img = acc[..., 0:3] / acc[..., 3:4]
    rimg = ref[..., 0:3] / ref[..., 3:4]
    assert img.mean() > 1e-3  # the lamp lights the scene
    diff = np.abs(img - rimg) / np.maximum(np.abs(rimg), 1e-3)
    assert np.mean(diff < 1e-4) > 0.97, np.mean(diff < 1e-4)
    assert np.mean((img - rimg) ** 2 / (rimg * rimg + 1e-2)) < 1e-5
