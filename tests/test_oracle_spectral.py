"""The oracle's hero-wavelength transport (oracle/spectral.c and the spectral
mode of oracle/pt.c and oracle/medium.c), pinned by what the mathematics
fixes; the mode has no counterpart in sprout, which renders RGB:

- the PCG hash against a numpy restatement of Jarzynski and Olano's
  published function, and the hero numbers it gives uniform (KS test);
- the four wavelengths: quarter-range apart, covering [380, 720) nm;
- the bands: a partition of unity, so grey stays grey;
- the film's response: the dual basis of the bands, checked against a
  fine numpy quadrature of the Gram matrix (E[response_i b_j] = delta_ij),
  so a spectrum in the bands' span reads back as its own rgb on average;
- four-wavelength ratio tracking: exp(-mu_t(lambda) d) per wavelength in a
  homogeneous chromatic medium;
- the spectral path tracer: grey scenes render the rgb image (same random
  numbers per path, so the images agree pixel by pixel up to the film's
  wavelength noise), and single-scattering transport, linear in the
  colours, renders the rgb image in expectation also for chromatic walls.
"""

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O


def np_pcg_hash(v):
    v = np.uint64(v) & np.uint64(0xFFFFFFFF)
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(0xFFFFFFFF)
    shift = (state >> np.uint64(28)) + np.uint64(4)
    word = (((state >> shift) ^ state) * np.uint64(277803737)) & np.uint64(0xFFFFFFFF)
    return int((word >> np.uint64(22)) ^ word)


def np_hero(seed, sequence):
    h = np_pcg_hash((seed & 0xFFFFFFFF) ^ np_pcg_hash(seed >> 32))
    h = np_pcg_hash(h ^ (sequence & 0xFFFFFFFF))
    h = np_pcg_hash(h ^ (sequence >> 32))
    return np.float32(h >> 8) * np.float32(2.0 ** -24)


@pytest.mark.parametrize("seed,seq", [(1, 0), (1, 5), (7, 2**33 + 9), (2**40 + 3, 123456789)])
def test_hero_number_is_the_published_hash(seed, seq):
    assert O.spectral_hero(seed, seq) == float(np_hero(seed, seq))


def test_hero_numbers_are_uniform_and_wavelengths_stratified():
    u = np.array([O.spectral_hero(3, s) for s in range(20000)])
    assert u.min() >= 0 and u.max() < 1
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    for x in (0.0, 0.1, 0.5, 0.74999, 0.9999):
        lam = sorted(O.spectral_wavelength(np.float32(x), k) for k in range(4))
        assert lam[0] >= 380 and lam[-1] < 720
        assert np.allclose(np.diff(lam), 85, atol=1e-3)


def test_bands_are_a_partition_of_unity():
    for lam in np.linspace(380, 719.99, 997):
        b = O.spectral_bands(lam)
        assert np.all(b >= 0) and np.all(b <= 1)
        assert abs(b.sum() - 1) < 1e-6
        x = np.float32(0.37)
        assert abs(O.spectral_upsample(b, [x, x, x]) - x) < 1e-7
    assert np.allclose(O.spectral_bands(400), [0, 0, 1])
    assert np.allclose(O.spectral_bands(530), [0, 1, 0])
    assert np.allclose(O.spectral_bands(700), [1, 0, 0])
    assert np.allclose(O.spectral_bands(480), [0, 0.5, 0.5], atol=1e-6)
    assert np.allclose(O.spectral_bands(585), [0.5, 0.5, 0], atol=1e-6)


def test_response_is_the_dual_basis():
    lam = 380 + (np.arange(340000) + 0.5) * 340 / 340000
    b = np.stack([O.spectral_bands(x) for x in lam[::100]])  # one in 100 through the oracle
    blue = np.clip((510 - lam) / 60, 0, 1)
    red = np.clip((lam - 560) / 50, 0, 1)
    bands = np.stack([red, 1 - blue - red, blue], 1)
    assert np.allclose(b, bands[::100], atol=1e-6)
    gram = bands.T @ bands / len(lam)
    minv = O.spectral_response().astype(np.float64)
    assert np.allclose(minv @ gram, np.eye(3), atol=1e-4)
    # the four-wavelength film estimate of a spectrum in the bands' span is
    # its rgb on average over the hero
    rgb = np.array([0.8, 0.3, 0.1], np.float32)
    us = (np.arange(4000) + 0.5) / 4000
    est = np.zeros(3)
    for u in us:
        b4 = np.stack([O.spectral_bands(O.spectral_wavelength(np.float32(u), k)) for k in range(4)])
        v4 = np.array([O.spectral_upsample(b, rgb) for b in b4], np.float32)
        est += O.spectral_to_rgb(b4, v4)
    assert np.allclose(est / len(us), rgb, atol=2e-3)


def test_four_wavelength_ratio_tracking_is_exact_per_wavelength():
    """a homogeneous chromatic medium of density 0.8: residual ratio
    tracking gives exp(-mu_t(lambda) d) for each of four channels with no
    random steps (minorant == majorant), and the delta-tracking weights of a
    scatter are mu_s / mu_t per channel in expectation."""
    grid = np.full((8, 8, 8), 0.8, np.float32)
    cc_a, cc_s = (0.3, 0.2, 0.1), (0.6, 0.9, 1.1)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
    lam = [420.0, 500.0, 590.0, 680.0]
    b4 = np.stack([O.spectral_bands(x) for x in lam])
    a4 = np.array([O.spectral_upsample(b, cc_a) for b in b4], np.float32)
    s4 = np.array([O.spectral_upsample(b, cc_s) for b in b4], np.float32)
    # the majorant (max over rgb) bounds every wavelength's coefficient
    assert np.all(a4 + s4 <= max(cc_a) + max(cc_s) + 1e-6)
    rays = np.zeros((4, 8), np.float32)
    rays[:, 0:3] = (0.5, -1.0, 0.5)
    rays[:, 4:7] = (0, 1, 0)
    rays[:, 7] = 1e30
    out = O.medium_channels(m, rays, a4, s4, transmittance=True)
    t0 = float(np.float32(O.offset_f(1.0)))
    # min over rgb of the coefficients is the minorant: the residual is
    # ratio tracked, so the estimate is exact only in the mean; use many rays
    n = 4000
    many = np.repeat(rays[:1], n, axis=0)
    tr = O.medium_channels(m, many, a4, s4, transmittance=True, seed=5)[:, 2:]
    expect = np.exp(-(2.0 - t0) * 0.8 * (a4 + s4))
    sem = tr.std(axis=0) / np.sqrt(n) + 1e-6
    assert np.all(np.abs(tr.mean(axis=0) - expect) < 5 * sem + 1e-4), (tr.mean(axis=0), expect)
    assert np.all(out[:, 0] == 1)
    # delta tracking: scatter probability and weights per channel
    ev = O.medium_channels(m, many, a4, s4, seed=6)
    sc = ev[:, 0] == 1
    d = 2.0 - t0
    for c in range(4):
        mu_t = 0.8 * (a4[c] + s4[c])
        want = (0.8 * s4[c]) / mu_t * (1 - np.exp(-mu_t * d))
        got = np.where(sc, ev[:, 2 + c], 0).mean()
        err = np.where(sc, ev[:, 2 + c], 0).std() / np.sqrt(n)
        assert abs(got - want) < 5 * err + 1e-3, (c, got, want)


def grey_room():
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    mats = mats.copy()
    view = mats.view(O.MATERIAL_DTYPE)
    for i in range(len(view)):
        c = view[i]["color"]
        g = np.float32(np.mean(c))
        view[i]["color"] = (g, g, g)
    return pos, idx, tmat, mats


def test_grey_room_spectral_equals_rgb():
    """grey walls and boxes (the light keeps its colour): the paths draw the
    same random numbers and see wavelength-independent reflectance, so the
    spectral and rgb images differ only by the film's wavelength weights,
    whose mean is one."""
    from scene_util import light_tree
    from sprout_amd import scenes

    pos, idx, tmat, mats = grey_room()
    tree = light_tree(pos, idx, tmat, mats)
    w, h, spp = 32, 32, 64
    rgb = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA, tree, seed=3) / spp
    spec = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA, tree, seed=3,
                    spectral=True) / spp
    assert np.all(np.isfinite(spec))
    assert abs(spec.mean() / rgb.mean() - 1) < 0.01
    blocks = lambda x: x.reshape(8, 4, 8, 4, 3).mean(axis=(1, 3))  # noqa: E731
    rel = np.abs(blocks(spec) - blocks(rgb)) / np.maximum(blocks(rgb), 1e-2)
    assert np.median(rel) < 0.02, np.median(rel)


def test_single_scattering_spectral_equals_rgb_in_expectation():
    """one bounce (max_bounces 1): the pixel is linear in the wall colours,
    so the spectral estimate of the chromatic room converges to the rgb
    image, colour by colour."""
    from scene_util import light_tree
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    tree = light_tree(pos, idx, tmat, mats)
    w, h, spp = 16, 16, 1024
    kw = dict(min_bounces=1, max_bounces=1)
    rgb = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA, tree, seed=4, **kw) / spp
    spec = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA, tree, seed=5,
                    spectral=True, **kw) / spp
    for c in range(3):
        a, b = spec[..., c].mean(), rgb[..., c].mean()
        assert abs(a / b - 1) < 0.02, (c, a, b)
    # and multiple scattering of the chromatic room does differ from rgb:
    # reflectance products are not linear in the colours
    red = np.float32(0.63), np.float32(0.06), np.float32(0.05)
    b4 = np.stack([O.spectral_bands(x) for x in (400.0, 500.0, 600.0, 700.0)])
    sq = np.array([O.spectral_upsample(b, red) ** 2 for b in b4], np.float32)
    assert not np.allclose(O.spectral_to_rgb(b4, sq), np.square(red), atol=1e-3)
