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
    The guide samples the first two vertices (fraction 0.8), trained on
    8192 spp of the 64 x 64 image; the comparison runs over eight 64 spp
    pairs against a 16384 spp reference."""
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    sc = scene(gpu, 0.05)
    cam = render.camera(64, 64, **scenes.ROOM_CAMERA)
    log("slit room reference, 64x64, 16384 spp")
    ref, _ = render.render(sc, cam, 16384, samples_per_wave=512, seed=31)
    ref = ref.cpu().numpy() / 16384
    e_plain, e_guided = [], []
    g = guide.Guide(5)
    log("training the guide, 8192 spp")
    render.render_guided(sc, cam, 8192, g, fraction=0.8, train=True, train_steps=4,
                         learning_rate=3e-3, samples_per_wave=64, seed=32, guide_depth=2,
                         record_depth=2)
    torch.cuda.synchronize()
    log("equal-spp comparison, 8 x 64 spp")
    for k in range(8):
        plain, _ = render.render(sc, cam, 64, samples_per_wave=64, seed=33 + 2 * k)
        guided, _ = render.render_guided(sc, cam, 64, g, fraction=0.8, train=False,
                                         samples_per_wave=64, seed=34 + 2 * k, guide_depth=2)
        e_plain.append(rel_mse(plain.cpu().numpy() / 64, ref))
        e_guided.append(rel_mse(guided.cpu().numpy() / 64, ref))
    g.release()
    ratio = float(np.mean(e_guided) / np.mean(e_plain))
    log(f"relMSE plain {np.mean(e_plain):.4g} guided {np.mean(e_guided):.4g} ratio {ratio:.3f}")
    assert ratio < 0.95, (e_plain, e_guided)
    sc.release()


def test_training_is_reproducible(gpu):
    """Two trainings from the same start on the same seeds give the same
    parameters bit for bit, graph-captured or launched kernel by kernel."""
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    sc = scene(gpu, 0.05)
    cam = render.camera(48, 40, **scenes.ROOM_CAMERA)
    params = []
    for graph in (True, True, False):
        g = guide.Guide(9)
        log(f"training, graph {graph}")
        render.render_guided(sc, cam, 64, g, fraction=0.5, train=True, train_steps=3, batch=2048,
                             learning_rate=2e-3, samples_per_wave=8, seed=41, record_depth=2,
                             graph=graph)
        torch.cuda.synchronize()
        params.append(g.params())
        g.release()
    fresh = guide.Guide(9)
    assert not np.array_equal(params[0], fresh.params())
    fresh.release()
    assert np.array_equal(params[0], params[1])
    assert np.array_equal(params[0], params[2])
    sc.release()


def test_guided_slit_room_1024_converges_to_the_reference(gpu):
    """The guided render of the slit room at 1024 x 1024 (trained online at
    that size, one record per camera path) against the 65536-spp FP64
    reference: its error at 16 spp is within the plain render's noise at 16
    spp (per pixel and over 16 x 16 blocks) and its mean matches the
    reference's, so the guided estimator converges to the same image."""
    import os

    import torch

    from sprout_amd import scenes

    here = os.path.dirname(os.path.abspath(__file__))
    with np.load(os.path.join(here, "golden", "slit_room_ref_1024.npz"), allow_pickle=False) as z:
        ref = z["image"].astype(np.float64)
    render, guide = gpu
    sc = scene(gpu, 0.05)
    cam = render.camera(1024, 1024, **scenes.ROOM_CAMERA)
    g = guide.Guide(7)
    log("slit room 1024x1024: training 4 waves of 1 spp")
    _, st = render.render_guided(sc, cam, 4, g, fraction=0.5, train=True, train_steps=4,
                                 batch=65536, learning_rate=2e-3, record_depth=1, guide_depth=1,
                                 samples_per_wave=1, seed=51, stats=True)
    assert st.trained > 0
    log("plain and guided, 16 spp each")
    plain, _ = render.render(sc, cam, 16, samples_per_wave=4, seed=52)
    guided, _ = render.render_guided(sc, cam, 16, g, fraction=0.5, train=False, guide_depth=1,
                                     samples_per_wave=4, seed=53)
    torch.cuda.synchronize()
    plain = plain.cpu().numpy() / 16
    guided = guided.cpu().numpy() / 16
    g.release()
    sc.release()
    assert np.all(np.isfinite(guided))

    def blocks16(img):
        return img.reshape(64, 16, 64, 16, 3).mean(axis=(1, 3))

    e_plain, e_guided = rel_mse(plain, ref), rel_mse(guided, ref)
    b_plain, b_guided = rel_mse(blocks16(plain), blocks16(ref)), rel_mse(blocks16(guided), blocks16(ref))
    log(f"relMSE plain {e_plain:.4g} guided {e_guided:.4g}; 16x16 blocks plain {b_plain:.4g} "
        f"guided {b_guided:.4g}; means {plain.mean():.5g} {guided.mean():.5g} ref {ref.mean():.5g}")
    assert e_guided < 1.25 * e_plain
    assert b_guided < 1.5 * b_plain
    # the reference is the same image: over 16 x 16 blocks the 16-spp plain
    # render is within its noise of it (a mirrored or shifted reference
    # leaves a bias of about 3e-3 here)
    assert b_plain < 3e-4
    assert abs(guided.mean() / ref.mean() - 1) < 1e-2
    assert abs(plain.mean() / ref.mean() - 1) < 1e-2


def test_guided_graph_equals_launches(gpu):
    """The captured guided wave replays the kernel-by-kernel run bit for bit,
    with a fixed guide and with on-device training (records in fixed slots)."""
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
        img, st = render.render_guided(sc, cam, 1024, g, fraction=0.5, train=True, train_steps=3,
                                       batch=4096, learning_rate=2e-3, samples_per_wave=128,
                                       seed=5, graph=graph, stats=True)
        torch.cuda.synchronize()
        out.append((img.cpu().numpy(), st.trained, g.params()))
        g.release()
    log(f"trained guided, graph vs launches: trained {out[0][1]} / {out[1][1]}")
    assert out[0][1] > 0 and out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][2], out[1][2])
    sc.release()


@pytest.fixture(scope="module")
def medium_room(gpu):
    """The box room holding a smooth random 40 x 48 x 56 cloud (density up to
    3, albedo 0.75-0.55, Henyey-Greenstein g 0.5) around its middle."""
    from scipy.ndimage import gaussian_filter

    from sprout_amd import medium as M

    render, _ = gpu
    rng = np.random.default_rng(1)
    grid = gaussian_filter(rng.random((40, 48, 56)), 2.0)
    grid = (grid - grid.min()) / (grid.max() - grid.min())
    grid[grid < 0.3] = 0.0
    grid = (grid * 3.0).astype(np.float32)
    sc = scene(gpu, 0.0)
    M.set_medium(sc, grid, (-0.6, 0.3, -0.6), (0.6, 1.5, 0.6), (0.3, 0.4, 0.5), (1.2, 0.9, 0.6),
                 0.5)
    yield sc
    sc.release()


@pytest.mark.parametrize("fraction,train,roulette", [(0.5, False, False), (1.0, False, True),
                                                    (0.8, True, True)])
def test_guided_medium_vertices_converge_to_plain(gpu, medium_room, fraction, train, roulette):
    """Scattering vertices in the medium sample the one-sample mixture of the
    guide and the phase function (guide_medium): with the mixture pdf in the
    path weight and in the MIS weights of their shadow rays the render,
    trained or not, converges to the plain MIS image without Russian
    roulette.  The plain render is the reference without roulette: sprout's
    roulette divides by a continuation probability above one when the
    throughput exceeds one, which a chromatic medium's tracking weights do
    now and then; guided paths cap that probability at one."""
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    g = guide.Guide(5)
    cam = render.camera(32, 32, **scenes.ROOM_CAMERA)
    log(f"guided medium vertices, fraction {fraction} train {train} roulette {roulette}: "
        "4096 spp vs plain")
    ref, _ = render.render(medium_room, cam, 4096, samples_per_wave=256, seed=31, min_bounces=8,
                           max_bounces=8)
    acc, st = render.render_guided(medium_room, cam, 4096, g, fraction=fraction, train=train,
                                   train_steps=2, samples_per_wave=256, seed=32, stats=True,
                                   guide_medium=True, min_bounces=4 if roulette else 8,
                                   max_bounces=8)
    torch.cuda.synchronize()
    a = blocks(acc.cpu().numpy() / 4096)
    b = blocks(ref.cpu().numpy() / 4096)
    e = rel_mse(a, b)
    log(f"  relMSE of 4x4 blocks {e:.3g}, means {a.mean():.5g} {b.mean():.5g}")
    assert np.all(np.isfinite(a))
    assert e < 2e-3, e
    assert abs(a.mean() / b.mean() - 1) < 1e-2
    if train:
        assert st.trained > 0
    g.release()


def test_guide_medium_switch(gpu, medium_room):
    """guide_medium changes the samples only where a path scatters in a
    medium: on the room without a medium the switch gives the same image bit
    for bit, with one it gives another image."""
    import torch

    from sprout_amd import scenes

    render, guide = gpu
    g = guide.Guide(6)
    cam = render.camera(32, 32, **scenes.ROOM_CAMERA)
    plain_room = scene(gpu, 0.0)
    out = {}
    for name, sc in (("room", plain_room), ("medium", medium_room)):
        for on in (False, True):
            img, _ = render.render_guided(sc, cam, 64, g, fraction=0.5, train=False,
                                          samples_per_wave=16, seed=7, guide_medium=on)
            torch.cuda.synchronize()
            out[name, on] = img.cpu().numpy()
    assert np.array_equal(out["room", False], out["room", True])
    assert not np.array_equal(out["medium", False], out["medium", True])
    plain_room.release()
    g.release()
