"""Moving emitters (mi_scene_update_positions): new vertex positions rebuild
every emissive triangle's light, the parts' Primitive_trees, the scene light
tree (the device builder) and the triangle tree (the GPU SBVH builder).

- The scene light tree and its light records after an update equal, bit for
  bit, those of a scene created from the new positions (whose trees the host
  builder makes), and so does a render on the same seeds; moving back gives
  the first image again.
- After a move, every pick's pdf equals Tree::pdf of the picked light (the
  reference's debug check, scene.cpp:158-168) and the pdfs over all lights
  sum to one.
- An update that would leave an emissive triangle without area is refused
  and the scene keeps its positions.
- The batched scene baked into one soup (1.32 M triangles, 10 000 emitters)
  moves its emitters with the same equality to a fresh scene.
"""

import numpy as np
import pytest

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


def lit_room():
    """The box room with three more emissive parts of 2, 40 and 700 small
    triangles (the last with leaves of up to 10 lights, so leaf
    distributions), each its own material."""
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    rng = np.random.default_rng(3)
    pos = [pos]
    idx = [idx]
    tm = [tmat]
    nv = len(pos[0])
    extra = np.zeros(3, mats.dtype)
    for k, (e, c) in enumerate([((3, 2, 1), (0.6, 1.6, -0.4)), ((0.5, 2, 4), (-0.5, 1.2, 0.3)),
                                ((1.5, 1.5, 1.5), (0.2, 0.8, 0.5))]):
        extra[k]["kind"] = 2  # MI_MATERIAL_LIGHT
        extra[k]["emission"] = e
        n = (2, 40, 700)[k]
        centers = np.asarray(c) + rng.uniform(-0.15, 0.15, (n, 3))
        v = (centers[:, None, :] + rng.uniform(-0.03, 0.03, (n, 3, 3))).reshape(-1, 3)
        pos.append(v.astype(np.float32))
        idx.append((nv + np.arange(3 * n)).reshape(-1, 3).astype(np.uint32))
        tm.append(np.full(n, len(mats) + k, np.uint32))
        nv += 3 * n
    return (np.concatenate(pos), np.concatenate(idx), np.concatenate(tm),
            np.concatenate([mats, extra]))


def moved(pos, idx, tmat, mats, seed):
    """the emitters moved: each emissive part translated and its vertices
    jittered"""
    rng = np.random.default_rng(seed)
    p = pos.copy()
    em = mats["kind"][tmat] == 2
    verts = idx[em].reshape(-1)
    owner = np.repeat(tmat[em], 3)
    v, first = np.unique(verts, return_index=True)
    shift = rng.uniform(-0.2, 0.2, (len(mats), 3)).astype(np.float32)
    p[v] += shift[owner[first]]
    p[v] += rng.uniform(-0.01, 0.01, (len(v), 3)).astype(np.float32)
    return p


def image(gpu, sc, spp=32, seed=5):
    import torch

    from sprout_amd import scenes

    cam = gpu.camera(48, 40, **scenes.ROOM_CAMERA)
    acc, _ = gpu.render(sc, cam, spp, samples_per_wave=8, seed=seed)
    torch.cuda.synchronize()
    return acc.cpu().numpy()


def same_lights(a, b):
    from sprout_amd import lights

    ta, tb = lights.scene_tree(a), lights.scene_tree(b)
    for x, y in zip(ta, tb):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def test_update_equals_a_fresh_scene(gpu):
    pos, idx, tmat, mats = lit_room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    first = image(gpu, sc)
    new = moved(pos, idx, tmat, mats, 1)
    log("room with 4 emissive parts: update against a fresh scene")
    ms = sc.update_positions(new)
    log(f"  light rebuild {ms:.3f} ms")
    fresh = gpu.Scene(new, idx, tmat, mats)
    same_lights(sc, fresh)
    after = image(gpu, sc)
    assert not np.array_equal(after, first)
    assert np.array_equal(after, image(gpu, fresh))
    fresh.release()
    sc.update_positions(pos)
    assert np.array_equal(image(gpu, sc), first)
    sc.release()


def test_pdfs_are_consistent_after_a_move(gpu):
    from sprout_amd import lights

    pos, idx, tmat, mats = lit_room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    sc.update_positions(moved(pos, idx, tmat, mats, 2))
    rng = np.random.default_rng(5)
    m = 1000
    p = rng.uniform([-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], (m, 3)).astype(np.float32)
    nrm = rng.normal(size=(m, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm.astype(np.float32)
    u = rng.random(m).astype(np.float32)
    for split in (False, True):
        picks, pdfs, counts = lights.sample(sc, p, nrm, u, split)
        for k in range(max(counts)):
            rows = counts > k
            got = lights.pdf(sc, p[rows], nrm[rows], picks[rows, k], split)
            assert np.allclose(got, pdfs[rows, k], rtol=1e-6, atol=0)
    nl = len(lights.scene_tree(sc)[4])
    total = np.zeros(m)
    for light in range(nl):
        total += lights.pdf(sc, p, nrm, np.full(m, light, np.uint32), False)
    ok = total > 0
    assert ok.mean() > 0.9
    assert np.allclose(total[ok], 1.0, rtol=1e-4)
    sc.release()


def test_degenerate_emitter_is_refused(gpu):
    from sprout_amd import _lib

    pos, idx, tmat, mats = lit_room()
    sc = gpu.Scene(pos, idx, tmat, mats)
    first = image(gpu, sc, spp=8)
    bad = moved(pos, idx, tmat, mats, 3)
    t = int(np.nonzero(tmat == len(mats) - 1)[0][0])
    bad[idx[t, 1]] = bad[idx[t, 0]]
    bad[idx[t, 2]] = bad[idx[t, 0]]  # an emissive triangle collapsed to a point
    with pytest.raises(_lib.MiError):
        sc.update_positions(bad)
    assert np.array_equal(image(gpu, sc, spp=8), first)
    with pytest.raises(_lib.MiError):
        sc.update_positions(pos[:-3])
    sc.release()


def test_batch_soup_moves_its_emitters(gpu):
    import time

    import torch

    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.batch_scene()
    log(f"batch soup: {len(idx)} triangles")
    sc = gpu.Scene(pos, idx, tmat, mats)
    new = moved(pos, idx, tmat, mats, 4)
    dev = torch.from_numpy(new).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = sc.update_positions(dev)
    torch.cuda.synchronize()
    log(f"  update {1e3 * (time.perf_counter() - t0):.1f} ms, of it lights {ms:.2f} ms")
    fresh = gpu.Scene(new, idx, tmat, mats)
    same_lights(sc, fresh)
    fresh.release()
    sc.release()
