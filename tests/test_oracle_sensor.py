"""Filtered sensor: the oracle's filters (oracle/sensor.c) and splat
(oracle.filtered_sums) pinned before they check the GPU.

- Gaussian, Blackman and Mitchell equal the reference's own filter classes,
  compiled here (oracle/_ref), on a dense grid of offsets;
- the splat of a constant colour resolves to that colour whatever the
  jitter (the weights cancel), and the box filter gives the plain per-pixel
  sums;
- a sample right of its pixel centre weighs more on the right neighbour (the
  sign of the offsets), and the clamps follow clamp.inl;
- the path tracer's border grid: with no border its samples sum to the
  per-pixel sums of the plain render.
"""

import numpy as np
import pytest

from oracle import oracle as O

CASES = [(O.FILTER_GAUSSIAN, 2.0, 1.8, 0.0, 0.0), (O.FILTER_GAUSSIAN, 1.0, 1.8, 0.0, 0.0),
         (O.FILTER_GAUSSIAN, 1.5, 2.5, 0.0, 0.0), (O.FILTER_BLACKMAN, 2.0, 0.0, 0.0, 0.0),
         (O.FILTER_BLACKMAN, 1.0, 0.0, 0.0, 0.0), (O.FILTER_MITCHELL, 2.0, 0.0, 1 / 3, 1 / 3),
         (O.FILTER_MITCHELL, 1.0, 0.0, 0.5, 0.25)]


@pytest.mark.parametrize("kind,radius,alpha,b,c", CASES)
def test_filters_equal_the_reference(kind, radius, alpha, b, c):
    l = O.ref_filter_lib()
    if l is None:
        pytest.skip("the reference library is not built here (oracle/ref.mk)")
    d = np.linspace(-2.6, 2.6, 4001).astype(np.float32)
    mine = O.Filter(kind, radius, alpha, b, c)(d)
    ref = O.ref_filter_eval(l, kind, radius, alpha, b, c, d)
    # the reference build contracts some products into FMAs (a few ulps)
    assert np.allclose(mine, ref, rtol=2e-6, atol=2e-6), np.max(np.abs(mine - ref))


def test_gaussian_closed_form():
    """Within the table's resolution the Gaussian is the normalized
    exp(-alpha d^2) - exp(-alpha r^2), clipped at r"""
    f = O.Filter(O.FILTER_GAUSSIAN, 2.0, 1.8)
    d = np.linspace(0, 1.9, 50).astype(np.float32)
    g = np.exp(-1.8 * d.astype(np.float64) ** 2) - np.exp(-1.8 * 4.0)
    x = np.linspace(0, 2, 200001)
    norm = 2 * np.trapezoid(np.exp(-1.8 * x * x) - np.exp(-1.8 * 4.0), x)
    assert np.allclose(f(d), g / norm, rtol=0, atol=5e-3)
    assert np.all(f(np.array([2.0, 2.5, -3.0], np.float32)) == 0)


def samples_of(colour, spp, h, w, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((spp, h, w, 5), np.float32)
    s[..., 0:3] = colour
    s[..., 3:5] = rng.random((spp, h, w, 2))
    return s


@pytest.mark.parametrize("kind,radius", [(O.FILTER_GAUSSIAN, 2.0), (O.FILTER_BLACKMAN, 2.0),
                                         (O.FILTER_MITCHELL, 2.0), (O.FILTER_GAUSSIAN, 1.0)])
def test_constant_colour_resolves_to_itself(kind, radius):
    f = O.Filter(kind, radius)
    bd = f.border()
    s = samples_of((0.25, 0.5, 2.0), 3, 9 + 2 * bd, 11 + 2 * bd, 1)
    acc = O.filtered_sums(s, 11, 9, f)
    img = acc[..., 0:3] / acc[..., 3:4]
    assert np.all(acc[..., 3] > 0)
    assert np.allclose(img, [0.25, 0.5, 2.0], rtol=1e-6)


def test_box_is_the_plain_sum():
    rng = np.random.default_rng(2)
    s = samples_of(0.0, 4, 6, 7, 3)
    s[..., 0:3] = rng.random((4, 6, 7, 3))
    acc = O.filtered_sums(s, 7, 6, O.Filter(O.FILTER_BOX))
    assert np.allclose(acc[..., 0:3], s[..., 0:3].astype(np.float64).sum(0))
    assert np.all(acc[..., 3] == 4)


def test_offsets_point_the_right_way():
    f = O.Filter(O.FILTER_GAUSSIAN, 2.0)
    s = np.zeros((1, 9, 9, 5), np.float32)
    s[0, 4, 4, 0:3] = 1.0
    s[0, 4, 4, 3] = 0.9  # right of the centre
    s[0, 4, 4, 4] = 0.5
    acc = O.filtered_sums(s, 5, 5, f)
    # grid pixel (4, 4) is image pixel (2, 2); offset +0.4: f(0.4) at the
    # centre, f(-0.6) on the right, f(1.4) on the left
    assert acc[2, 2, 0] > acc[2, 3, 0] > acc[2, 1, 0] > 0
    assert np.isclose(acc[2, 3, 0], acc[2, 2, 0] * f(np.float32(-0.6))[0] / f(np.float32(0.4))[0])
    assert np.isclose(acc[1, 2, 0], acc[3, 2, 0])


def test_clamps():
    c = np.array([[2.0, 0.5, 4.0]], np.float32)
    assert np.allclose(O.clamp_colors(c, (1.0, 1.0, 1.0)), [[1.0, 0.5, 1.0]])
    assert np.allclose(O.clamp_colors(c, (2.0, -1.0, 0.0)), [[1.0, 0.25, 2.0]])
    assert np.array_equal(O.clamp_colors(c, (-1.0, 0.0, 0.0)), c)


def test_border_grid_samples_sum_to_the_plain_render():
    from scene_util import light_tree
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    tree = light_tree(pos, idx, tmat, mats)
    plain = O.render(pos, idx, tmat, mats, 12, 10, 3, scenes.ROOM_CAMERA, tree, seed=4)
    s = O.render(pos, idx, tmat, mats, 12, 10, 3, scenes.ROOM_CAMERA, tree, seed=4, samples=True)
    assert s.shape == (3, 10, 12, 5)
    assert np.allclose(s[..., 0:3].astype(np.float64).sum(0), plain, rtol=1e-6, atol=1e-7)
    # with a border the grid grows and the image pixels see the same scene
    b = O.render(pos, idx, tmat, mats, 12, 10, 64, scenes.ROOM_CAMERA, tree, seed=5, border=2,
                 samples=True)
    assert b.shape == (64, 14, 16, 5)
    inner = b[:, 2:-2, 2:-2, 0:3].astype(np.float64)
    ref = O.render(pos, idx, tmat, mats, 12, 10, 64, scenes.ROOM_CAMERA, tree, seed=6,
                   samples=True)[..., 0:3].astype(np.float64)
    # the same image up to noise: the means agree within 4 standard errors
    se = np.sqrt(inner.var() / inner.size + ref.var() / ref.size)
    assert abs(inner.mean() - ref.mean()) < 4 * se
