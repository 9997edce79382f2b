"""The oracle's path tracer against the reference's own renderer.

tests/golden/sprout_renders.npz holds the takes of tests/golden/take_fixture
and tests/golden/take_smooth rendered by the reference's whole renderer
(take and scene loaders, rendering::Driver, Pathtracer_MIS, its sensor),
compiled here from its sources (oracle/ref_render.mk) and driven by
tests/golden/make_sprout_renders.py at 65 536 samples per pixel, without a
sensor filter.  The oracle renders the same takes, loaded by sprout_amd.take,
with other random numbers, so the images agree in expectation:

* small: a floor and a metallic triangle sharing two of its vertices, lit by
  a quad lamp.  The JSON mesh has no normals, so sprout shades with the last
  geometric normal written to each vertex (triangle_mesh_provider.cpp:
  103-118), which darkens the floor along the shared edge;
* smooth: a 12 x 6 UV sphere with explicit vertex normals in two parts
  (copper metal, blue dielectric) on a floor;
* room: the path-tracer benchmark's box room (its walls as Substitutes of
  roughness 1, its ceiling light) from the benchmark's camera, 64 x 64.

Each check is also run with a planted mistake, which it must fail: the
geometric normal instead of the shading normal (small, smooth), the colours
without sprout's sRGB -> ACEScg mapping (room).
Measured at 4096 spp: image means within 0.1 %, 16 x 16-pixel blocks within
0.4 % (small), 1.3 % (smooth) and 0.2 % (room).  The smooth take's residual
sits on the sphere's silhouette in shadow, where sprout's one-bounce indirect light is
lower than the oracle's (whose MIS, BSDF-only and light-only estimators agree
there); with flat normals the same pixels agree.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O
from scene_util import light_tree

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
TAKES = {"small": os.path.join(GOLDEN, "take_fixture", "small.take"),
         "smooth": os.path.join(GOLDEN, "take_smooth", "smooth.take"),
         "room": os.path.join(GOLDEN, "take_room", "room.take")}
SPP = 4096
# (image mean, 16 x 16 block) relative tolerances per take
TOLERANCE = {"small": (0.005, 0.01), "smooth": (0.005, 0.025), "room": (0.005, 0.01)}


@pytest.fixture(scope="module")
def sprout_images():
    with np.load(os.path.join(GOLDEN, "sprout_renders.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def blocks(img, b):
    h, w, _ = img.shape
    return img[:h // b * b, :w // b * b].reshape(h // b, b, w // b, b, 3).mean(axis=(1, 3))


def errors(img, ref):
    mean = np.abs(img.mean(axis=(0, 1)) / ref.mean(axis=(0, 1)) - 1).max()
    rb, ib = blocks(ref, 16), blocks(img, 16)
    block = (np.abs(ib - rb) / np.maximum(rb, 2e-3)).max()
    return mean, block


def oracle_image(name, normals=True, srgb=False):
    """srgb: the scene's colours as written, not mapped to ACEScg"""
    import json

    from sprout_amd import take

    t = take.Take(TAKES[name])
    pos, idx, tmat, mats = t.scene
    if srgb:
        mats = mats.copy()
        with open(os.path.join(os.path.dirname(TAKES[name]), "room.scene")) as f:
            doc = json.load(f)
        for k, m in enumerate(doc["materials"]):
            kind, node = next(iter(m["rendering"].items()))
            key = "emission" if kind == "Light" else "color"
            mats[key][k] = node[key]
    acc = O.render(pos, idx, tmat, mats, t.width, t.height, SPP, t.camera,
                   light_tree(pos, idx, tmat, mats), seed=7, min_bounces=t.min_bounces,
                   max_bounces=t.max_bounces, normals=t.normals if normals else None)
    return acc / SPP


@pytest.mark.parametrize("name", sorted(TAKES))
def test_oracle_converges_to_sprouts_render(sprout_images, name):
    ref = sprout_images[name]
    assert ref.shape[2] == 3 and int(sprout_images["spp"]) == 65536
    mean, block = errors(oracle_image(name), ref)
    tol_mean, tol_block = TOLERANCE[name]
    assert mean < tol_mean, mean
    assert block < tol_block, block


@pytest.mark.parametrize("name", sorted(TAKES))
def test_planted_mistakes_fail_the_check(sprout_images, name):
    """shading with the geometric normal (small, smooth), colours without
    the ACEScg mapping (room, whose flat walls need no shading normals)"""
    img = oracle_image(name, srgb=True) if name == "room" else oracle_image(name, normals=False)
    mean, block = errors(img, sprout_images[name])
    tol_mean, tol_block = TOLERANCE[name]
    assert mean > 4 * tol_mean or block > 4 * tol_block, (mean, block)


def test_filtered_sensor_converges_to_sprouts_render(sprout_images):
    """the fixture take through its Gaussian sensor (radius 2, alpha 1.8):
    samples over the filter-extended grid splatted by the oracle's filter,
    against the reference's Filtered_2p0 sensor.  The interior agrees to
    0.01 %; the two rows under the top edge, where the lamp's glow comes
    partly from samples beyond the image, are 1-2 % brighter in sprout's
    (0.4 % of the image mean)."""
    from sprout_amd import take

    t = take.Take(TAKES["small"])
    pos, idx, tmat, mats = t.scene
    s = O.render(pos, idx, tmat, mats, t.width, t.height, 512, t.camera,
                 light_tree(pos, idx, tmat, mats), seed=7, min_bounces=t.min_bounces,
                 max_bounces=t.max_bounces, border=2, samples=True, normals=t.normals)
    acc = O.filtered_sums(s, t.width, t.height, O.Filter(O.FILTER_GAUSSIAN, 2.0))
    img = acc[..., :3] / acc[..., 3:4]
    ref = sprout_images["small_filtered"]
    mean, block = errors(img, ref)
    assert mean < 0.006, mean
    assert block < 0.02, block
    inner = img[4:-4, 4:-4].mean() / ref[4:-4, 4:-4].mean() - 1
    assert abs(inner) < 0.003, inner
