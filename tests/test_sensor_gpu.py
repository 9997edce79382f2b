"""The filtered sensor (mi_render_filtered) against the oracle: the path
tracer's samples over the border-extended grid (oracle/pt.c) splatted by the
oracle's restatement of Filtered_1p0 / Filtered_2p0 with sprout's filters
(oracle/sensor.c, equal to the reference's filter classes by test).  Same
seeds give the same samples up to the device's sin/cos ulps; the box sensor
equals the plain render bit for bit."""

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


@pytest.fixture(scope="module")
def room(gpu):
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    yield sc, (pos, idx, tmat, mats)
    sc.release()


def test_box_sensor_is_the_plain_render(gpu, room):
    import torch

    from sprout_amd import _lib, scenes

    sc, _ = room
    cam = gpu.camera(40, 30, **scenes.ROOM_CAMERA)
    log("box sensor vs plain render, 40x30, 8 spp")
    a, _ = gpu.render_filtered(sc, cam, 8, gpu.sensor(_lib.MI_FILTER_BOX), samples_per_wave=4, seed=3)
    b, _ = gpu.render(sc, cam, 8, samples_per_wave=4, seed=3)
    torch.cuda.synchronize()
    a = a.cpu().numpy()
    assert np.array_equal(a[..., 0:3], b.cpu().numpy())
    assert np.all(a[..., 3] == 8)


@pytest.mark.parametrize("kind,radius,clamp", [
    (1, 2.0, (-1.0, 0.0, 0.0)), (1, 1.0, (-1.0, 0.0, 0.0)), (2, 2.0, (0.5, 0.5, 0.5)),
    (3, 2.0, (0.8, -1.0, 0.0)), (3, 1.5, (-1.0, 0.0, 0.0))])
def test_filtered_same_seeds_as_oracle(gpu, room, kind, radius, clamp):
    import torch

    from sprout_amd import scenes

    sc, (pos, idx, tmat, mats) = room
    w, h, spp = 40, 30, 8
    log(f"filter {kind} radius {radius} clamp {clamp}: {w}x{h}, {spp} spp")
    cam = gpu.camera(w, h, **scenes.ROOM_CAMERA)
    acc, _ = gpu.render_filtered(sc, cam, spp, gpu.sensor(kind, radius, clamp_max=clamp),
                                 samples_per_wave=4, seed=7)
    torch.cuda.synchronize()
    acc = acc.cpu().numpy()
    f = O.Filter(kind, radius)
    s = O.render(pos, idx, tmat, mats, w, h, spp, scenes.ROOM_CAMERA,
                 light_tree(pos, idx, tmat, mats), seed=7, border=f.border(), samples=True)
    ref = O.filtered_sums(s, w, h, f, clamp)
    assert np.all(np.isfinite(acc))
    # the weights: same tables, same jitter
    assert np.allclose(acc[..., 3], ref[..., 3], rtol=1e-5, atol=1e-6)
    img = acc[..., 0:3] / acc[..., 3:4]
    rimg = ref[..., 0:3] / ref[..., 3:4]
    diff = np.abs(img - rimg) / np.maximum(np.abs(rimg), 1e-3)
    assert np.mean(diff < 1e-4) > 0.97, np.mean(diff < 1e-4)
    assert np.mean((img - rimg) ** 2 / (rimg * rimg + 1e-2)) < 1e-5


def test_bad_sensors_are_rejected(gpu, room):
    from sprout_amd import _lib, scenes

    sc, _ = room
    cam = gpu.camera(8, 8, **scenes.ROOM_CAMERA)
    for bad in (gpu.sensor(1, 3.0), gpu.sensor(1, 0.0), gpu.sensor(7, 2.0), gpu.sensor(1, 2.0, alpha=-1.0)):
        with pytest.raises(_lib.MiError):
            gpu.render_filtered(sc, cam, 1, bad)
