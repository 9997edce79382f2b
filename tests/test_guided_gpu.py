"""Guided path tracing (mi_render_guided) on the GPU.

The guide only changes how directions are sampled; with the one-sample
mixture pdf in the path weight and in the MIS weights the estimator stays
unbiased, so guided renders, trained or not, converge to the plain MIS
image.  Training on the slit room (light behind a wall with a 0.05-wide
slit, most light indirect) must lower the error per sample.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh, guide, render

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield render, guide
    torch.cuda.synchronize()


def log(msg):
    print(msg, flush=True)


def rel_mse(x, r, eps=1e-2):
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def scene(gpu, slit):
    from sprout_amd import scenes

    render, _ = gpu
    pos, idx, tmat, mats = scenes.room(slit)
    return render.Scene(pos, idx, tmat, mats)


def blocks(img, k=4):
    h, w, _ = img.shape
    return img.reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))


@pytest.mark.parametrize("fraction,train", [(0.5, False), (0.5, True), (0.0, False)])
def test_guided_converges_to_plain(gpu, fraction, train):
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    sc = scene(gpu, 0.0)
    g = guide.Guide(3)
    cam = render.camera(32, 32, **scenes.ROOM_CAMERA)
    log(f"guided fraction {fraction} train {train}: 4096 spp vs plain 4096 spp")
    ref, _ = render.render(sc, cam, 4096, samples_per_wave=256, seed=21)
    acc, st = render.render_guided(sc, cam, 4096, g, fraction=fraction, train=train,
                                   train_steps=2, samples_per_wave=256, seed=22, stats=True)
    torch.cuda.synchronize()
    a = blocks(acc.cpu().numpy() / 4096)
    b = blocks(ref.cpu().numpy() / 4096)
    assert np.all(np.isfinite(a))
    assert rel_mse(a, b) < 2e-3, rel_mse(a, b)
    assert abs(a.mean() / b.mean() - 1) < 1e-2
    if train:
        assert st.trained > 0
    g.release()
    sc.release()


def test_training_lowers_error_in_the_slit_room(gpu):
    """Equal sample count after training: guided per-pixel error below plain.
    Training is not bitwise reproducible (records arrive in queue order), so
    the gain varies with the training: the mean ratio over four 64 spp pairs
    measured 0.84-0.97 over ten trainings on MI355X (centre 0.89).  Two
    trainings, four pairs each, are held to a mean ratio of 0.97."""
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    sc = scene(gpu, 0.05)
    cam = render.camera(64, 64, **scenes.ROOM_CAMERA)
    log("slit room reference, 64x64, 16384 spp")
    ref, _ = render.render(sc, cam, 16384, samples_per_wave=512, seed=31)
    ref = ref.cpu().numpy() / 16384
    e_plain, e_guided = [], []
    for trial in range(2):
        g = guide.Guide(5 + trial)
        log(f"training guide {trial}")
        render.render_guided(sc, cam, 512, g, fraction=0.5, train=True, train_steps=8,
                             learning_rate=2e-3, samples_per_wave=16, seed=32 + 100 * trial)
        torch.cuda.synchronize()
        log("equal-spp comparison, 4 x 64 spp")
        for k in range(4):
            plain, _ = render.render(sc, cam, 64, samples_per_wave=64, seed=33 + 2 * k)
            guided, _ = render.render_guided(sc, cam, 64, g, fraction=0.5, train=False,
                                             samples_per_wave=64, seed=34 + 2 * k)
            e_plain.append(rel_mse(plain.cpu().numpy() / 64, ref))
            e_guided.append(rel_mse(guided.cpu().numpy() / 64, ref))
        g.release()
    ratio = float(np.mean(e_guided) / np.mean(e_plain))
    log(f"relMSE plain {np.mean(e_plain):.4g} guided {np.mean(e_guided):.4g} ratio {ratio:.3f}")
    assert ratio < 0.97, (e_plain, e_guided)
    sc.release()


def test_guided_graph_equals_launches(gpu):
    """The captured guided wave replays the kernel-by-kernel run: bit for bit
    with a fixed guide; with on-device training both train and give the same
    image up to noise.  (Training records are appended in queue order, which
    varies from run to run, so trained guides are not bitwise reproducible.)"""
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    sc = scene(gpu, 0.05)
    cam = render.camera(32, 32, **scenes.ROOM_CAMERA)
    g = guide.Guide(3)
    render.render_guided(sc, cam, 64, g, fraction=0.5, train=True, train_steps=3, batch=4096,
                         samples_per_wave=8, seed=4)
    out = []
    for graph in (True, False):
        img, _ = render.render_guided(sc, cam, 32, g, fraction=0.5, train=False,
                                      samples_per_wave=8, seed=5, graph=graph)
        torch.cuda.synchronize()
        out.append(img.cpu().numpy())
    assert np.array_equal(out[0], out[1])
    g.release()

    out = []
    for graph in (True, False):
        g = guide.Guide(3)
        img, st = render.render_guided(sc, cam, 4096, g, fraction=0.5, train=True, train_steps=3,
                                       batch=4096, learning_rate=2e-3, samples_per_wave=128,
                                       seed=5 + graph, graph=graph, stats=True)
        torch.cuda.synchronize()
        out.append((blocks(img.cpu().numpy() / 4096), st.trained))
        g.release()
    e = rel_mse(out[0][0], out[1][0])
    log(f"trained guided, graph vs launches: relMSE {e:.3g}, trained {out[0][1]} / {out[1][1]}")
    assert out[0][1] > 0 and out[1][1] > 0
    assert e < 2e-2
    assert abs(out[0][0].mean() / out[1][0].mean() - 1) < 5e-2
    sc.release()
