"""The oracle's camera samples and perspective camera (oracle/sampler.c), and
the take loader's colour mapping, pinned against the reference's own
generator, Golden_ratio sampler and sRGB -> AP1 function compiled here
(oracle/ref.mk), and against what the mathematics fixes."""

import math

import numpy as np
import pytest

from oracle import oracle as O

REF = O.ref_sampler_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="the reference is not built here")


@needs_ref
@pytest.mark.parametrize("seed,seq", [(0, 0), (0, 1), (0, 123456789), (7, 2**40 + 3), (2**63, 5)])
def test_pcg_equals_the_reference_generator(seed, seq):
    n = 64
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    REF.sref_pcg_uint(seed, seq, n, u.ctypes.data)
    REF.sref_pcg_float(seed, seq, n, f.ctypes.data)
    assert np.array_equal(O.pcg_uint(seed, seq, n), u)
    assert np.array_equal(O.pcg_float(seed, seq, n).view(np.uint32), f.view(np.uint32))


def ref_samples(spp, seq, x=0, y=0):
    image = np.zeros((spp, 2), np.float32)
    lens = np.zeros((spp, 2), np.float32)
    time = np.zeros(spp, np.float32)
    REF.sref_golden_ratio_camera_samples(spp, seq, x, y, image.ctypes.data, lens.ctypes.data,
                                         time.ctypes.data)
    return image, lens, time


@needs_ref
@pytest.mark.parametrize("spp", [1, 2, 3, 4, 5, 7, 8, 13, 16, 21, 33, 64, 100, 256, 1000])
def test_golden_ratio_samples_equal_the_reference_sampler(spp):
    """Sampler::camera_sample of a Golden_ratio sampler started per pixel as
    Worker::render starts it: image, lens and time dimensions bit for bit."""
    rng = np.random.default_rng(spp)
    for seq in [0, 1, 65535, *rng.integers(0, 2**40, 5).tolist()]:
        want = ref_samples(spp, seq, 3, 4)
        got = O.golden_ratio_camera_samples(spp, seq)
        for w, g in zip(want, got):
            assert np.array_equal(w.view(np.uint32), g.view(np.uint32)), (spp, seq)


def golden_sequence(r, n):
    """r, r + phi, ... wrapped at 1, accumulated in float32"""
    out = np.zeros(n, np.float32)
    x = np.float32(r)
    for i in range(n):
        out[i] = x
        x = np.float32(x + np.float32(0.618033988749894))
        if x >= 1:
            x = np.float32(x - np.float32(1))
    return out


def test_golden_ratio_samples_are_two_golden_sequences():
    """Each axis of a pixel's samples holds exactly the points of one
    golden-ratio sequence (so the Fibonacci walk that orders the first axis
    visits every point once, and the shuffle only permutes), started at the
    pixel's two draws; the gaps between sorted points take at most three
    values (the three-distance theorem)."""
    for spp in (1, 2, 8, 13, 21, 34, 55, 89, 100, 256):
        seq = 12345 + spp
        img, _, _ = O.golden_ratio_camera_samples(spp, seq)
        first, second = O.pcg_float(0, seq, 2)
        for axis, start in ((0, second), (1, first)):
            want = np.sort(golden_sequence(start, spp))
            assert np.array_equal(np.sort(img[:, axis]), want), (spp, axis)
            v = np.sort(img[:, axis].astype(np.float64))
            gaps = np.diff(np.concatenate([v, [v[0] + 1]]))
            assert len(np.unique(np.round(gaps, 5))) <= 3
        # shuffled: not the generation order (for more than a handful)
        if spp >= 8:
            assert not np.array_equal(img[:, 1], golden_sequence(first, spp))


def test_pixel_sequence_numbers_the_border_grid():
    # rendering_worker.cpp:97-104: o0 = iteration * W'H', (y + r) W' + x + r
    assert O.pixel_sequence(0, 0, 10, 8, 0, 0) == 0
    assert O.pixel_sequence(3, 2, 10, 8, 0, 0) == 23
    assert O.pixel_sequence(-2, -2, 14, 12, 2, 0) == 0
    assert O.pixel_sequence(0, 0, 14, 12, 2, 1) == 14 * 12 + 2 * 14 + 2


def test_perspective_ray_closed_forms():
    """The centre of the image looks along r2, the left image edge at the
    horizontal half-angle toward -r0 (no mirror), the top edge toward +r1; the
    rays are unit length and leave from the camera position."""
    w, h, fov = 64, 48, 70.0
    rot = np.eye(3, dtype=np.float32).reshape(9)
    cam = {"position": [1.0, 2.0, 3.0], "rotation": rot, "fov_deg": fov}
    pixels = np.array([[32, 24], [0, 24], [32, 0], [63, 47]])
    uv = np.array([[0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [1.0, 1.0]], np.float32)
    r = O.camera_rays(w, h, cam, pixels, uv)
    assert np.allclose(r[:, 0:3], [1, 2, 3])
    d = r[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.allclose(d[0], [0, 0, 1], atol=1e-7)
    half = math.radians(fov / 2)
    assert d[1][0] < 0 and math.isclose(math.atan2(-d[1][0], d[1][2]), half, rel_tol=1e-6)
    t = math.tan(half) * h / w
    assert d[2][1] > 0 and math.isclose(math.atan2(d[2][1], d[2][2]), math.atan(t), rel_tol=1e-6)
    assert d[3][0] > 0 and d[3][1] < 0  # bottom right
    # a rotated camera: rows r0, r1, r2 map the camera axes
    a = math.radians(30)
    rot = np.array([[math.cos(a), 0, -math.sin(a)], [0, 1, 0], [math.sin(a), 0, math.cos(a)]],
                   np.float32)
    r2 = O.camera_rays(w, h, {"position": [0, 0, 0], "rotation": rot.reshape(9), "fov_deg": fov},
                       pixels, uv)
    assert np.allclose(r2[:, 4:7], d @ rot.astype(np.float64), atol=2e-7)


def test_look_at_agrees_with_the_renderer():
    """oracle.look_at and sprout_amd.render.look_at give the same camera."""
    from sprout_amd import render, scenes

    for w, h, c in ((256, 256, scenes.ROOM_CAMERA), (1920, 1080, scenes.BATCH_CAMERA),
                    (64, 48, scenes.CLOUD_CAMERA)):
        a = O.look_at(w, h, c["eye"], c["target"], c["up"], c["fov_deg"])
        b = render.look_at(w, h, c["eye"], c["target"], c["up"], c["fov_deg"])
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x, np.float32), np.asarray(y, np.float32))


@needs_ref
def test_srgb_to_ap1_equals_the_reference():
    from sprout_amd import take

    rng = np.random.default_rng(5)
    cols = np.concatenate([rng.random((200, 3)), rng.random((50, 3)) * 20,
                           [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [17, 12, 4]]]).astype(np.float32)
    want = np.zeros_like(cols)
    REF.sref_srgb_to_ap1(cols.ctypes.data, len(cols), want.ctypes.data)
    got = np.array([take.srgb_to_ap1(c) for c in cols])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_srgb_to_ap1_closed_forms():
    """White stays white (the matrix rows sum to 1 within its 8 digits), the
    mapping is linear, and grey numbers take the mapped green channel."""
    from sprout_amd import take

    assert np.allclose(take.srgb_to_ap1([1, 1, 1]), 1, atol=1e-6)
    a, b = np.array([0.2, 0.5, 0.1]), np.array([0.7, 0.1, 0.3])
    assert np.allclose(take.srgb_to_ap1(a + b), take.srgb_to_ap1(a) + take.srgb_to_ap1(b), atol=1e-6)
    g = take.srgb_to_ap1([0.3, 0.3, 0.3])[1]
    assert np.array_equal(take.read_color(0.3), np.array([g, g, g], np.float32))
    red = take.read_color([1, 0, 0])
    assert red[0] > 0.6 and red[1] > 0.06 and red[2] > 0.02  # AP1 red holds some green and blue
