"""The device light-tree builder (mi_light_tree_build_device,
mi_light_part_trees_build_device) against the host builder: the same trees
bit for bit — every node's bytes, the middles, the mapping and the orders,
and for Primitive_trees the leaf distributions — on random light sets from
one light to ten thousand (sweeps, the 16-slice grid, zero powers,
two-sided lights, opposite cones), on the batched scene's 10 065 emissive
parts, and on many parts built at once.  A node that no candidate splits
is reported (the host builder splits it at the median)."""

import numpy as np
import pytest

from sprout_amd import lights

pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield
    torch.cuda.synchronize()


def log(msg):
    print(msg, flush=True)


def random_lights(n, seed, zero_every=0, two_sided=False):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-2, 2, (n, 3)).astype(F)
    size = rng.uniform(0.01, 0.2, (n, 3)).astype(F)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cones = np.concatenate([nrm, rng.uniform(0.2, 1.0, (n, 1))], 1).astype(F)
    powers = np.exp(rng.normal(size=n)).astype(F)
    if zero_every:
        powers[::zero_every] = 0
    ts = (rng.random(n) < 0.3).astype(np.uint8) if two_sided else np.zeros(n, np.uint8)
    return (centers - size).astype(F), (centers + size).astype(F), cones, powers, ts


def same_tree(a, b):
    na, ma, pa, oa = a
    nb, mb, pb, ob = b
    assert len(na) == len(nb)
    assert np.array_equal(na.view(np.uint32), nb.view(np.uint32)), \
        np.nonzero(na.view(np.uint32) != nb.view(np.uint32))
    assert np.array_equal(ma, mb)
    assert np.array_equal(pa, pb)
    assert np.array_equal(oa, ob)


@pytest.mark.parametrize("n,seed,zero,two", [(1, 0, 0, False), (2, 1, 0, False), (5, 2, 0, True),
                                             (40, 3, 7, False), (127, 4, 0, True),
                                             (300, 5, 11, True), (2000, 6, 0, False),
                                             (10000, 7, 13, True)])
def test_scene_tree_equals_the_host_builder(gpu, n, seed, zero, two):
    lo, hi, cones, powers, ts = random_lights(n, seed, zero, two)
    log(f"scene tree of {n} lights")
    host = lights.build_tree(lo, hi, cones, powers, ts)
    dev = lights.build_tree_device(lo, hi, cones, powers, ts)
    same_tree(host, dev)


def test_opposite_cones_and_shared_points(gpu):
    """cones facing each other (the whole-sphere merge) and lights sharing
    box faces (equal candidate planes, equal costs: the first wins)"""
    n = 64
    lo, hi, cones, powers, ts = random_lights(n, 9)
    cones[::2, :3] = (0, 0, 1)
    cones[1::2, :3] = (0, 0, -1)
    hi[::3, 0] = hi[0, 0]
    hi[1::4, 2] = hi[1, 2]
    lo[::3, 0] = np.minimum(lo[::3, 0], hi[0, 0] - F(0.05))
    same_tree(lights.build_tree(lo, hi, cones, powers, ts),
              lights.build_tree_device(lo, hi, cones, powers, ts))


def test_batch_scene_parts_tree(gpu):
    """the scene tree over the batched scene's 10 065 emissive parts (10 000
    lights of log-normal power and the 65 others) as the scene keeps it"""
    from sprout_amd import render, scenes

    pos, idx, tmat, mats = scenes.batch_scene()
    sc = render.Scene(pos, idx, tmat, mats)
    nodes, middles, mapping, orders, recs = lights.scene_tree(sc)
    sc.release()
    n = len(recs)
    log(f"batch scene: {n} parts")
    c = recs["center"].astype(F)
    # the part boxes are not kept; rebuild the tree from boxes around the
    # centres and compare host and device on them
    r = np.asarray(recs["radius"], F)[:, None]
    lo, hi = (c - r).astype(F), (c + r).astype(F)
    ts = np.asarray(recs["two_sided"], np.uint8)
    same_tree(lights.build_tree(lo, hi, recs["cone"], recs["power"], ts),
              lights.build_tree_device(lo, hi, recs["cone"], recs["power"], ts))


def test_part_trees_equal_the_host_builder(gpu):
    sizes = [1, 3, 4, 5, 31, 32, 33, 200, 700, 5000]
    rng = np.random.default_rng(11)
    parts = [random_lights(s, 20 + k, zero_every=9 if k % 3 == 0 else 0) for k, s in enumerate(sizes)]
    first = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    lo = np.concatenate([p[0] for p in parts])
    hi = np.concatenate([p[1] for p in parts])
    cones = np.concatenate([p[2] for p in parts])
    powers = np.concatenate([p[3] for p in parts])
    pcs = []
    for p in parts:
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        pcs.append([*ax, rng.uniform(-0.5, 1.0)])
    pcs = np.array(pcs, F)
    log(f"{len(sizes)} part trees at once, {int(first[-1])} lights")
    dev = lights.build_part_trees_device(lo, hi, cones, powers, first, pcs)
    for k, s in enumerate(sizes):
        b, e = first[k], first[k + 1]
        host = lights.build_part_tree(lo[b:e], hi[b:e], cones[b:e], powers[b:e], pcs[k])
        same_tree(host[:4], dev[k][:4])
        assert np.array_equal(host[4], dev[k][4])
        assert np.array_equal(host[5].view(np.uint32), dev[k][5].view(np.uint32))
        if s > 300:
            assert (host[4] != 0xFFFFFFFF).any(), "no leaf distribution was exercised"


def test_unsplittable_node_is_reported(gpu):
    from sprout_amd import _lib

    n = 200
    lo, hi, cones, powers, ts = random_lights(n, 12)
    lo[:] = lo[0]
    hi[:] = hi[0]  # every light in one box: no candidate separates them
    with pytest.raises(_lib.MiError):
        lights.build_tree_device(lo, hi, cones, powers, ts)
    # the host builder splits at the median
    nodes, _, mapping, _ = lights.build_tree(lo, hi, cones, powers, ts)
    assert len(nodes) > 1 and sorted(mapping.tolist()) == list(range(n))
