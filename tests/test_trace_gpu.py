"""GPU parity of the traversal kernels (through the C ABI) against the oracle.

Trees are either the reference's own (golden fixture) or oracle-built, and
are uploaded with mi_bvh_upload; every result word (t, u, v, triangle) must
equal the pinned-reciprocal oracle bit for bit.
"""

import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NULL = 0xFFFFFFFF


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield bvh
    torch.cuda.synchronize()


def log(msg):
    print(msg, flush=True)


def to_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run_closest(gpu, tree, rays):
    import torch

    hits = tree.intersect(to_dev(rays.astype(np.float32)))
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(np.uint32)


def run_any(gpu, tree, rays):
    import torch

    occ = tree.intersect_p(to_dev(rays.astype(np.float32)))
    torch.cuda.synchronize()
    return occ.cpu().numpy().view(np.uint32)


def golden_upload(gpu, g):
    nodes = np.ascontiguousarray(g["nodes"]).view(O.NODE_DTYPE).reshape(-1)
    return gpu.upload(nodes, g["slots"], g["positions"], g["indices"]), O.Tree(nodes, g["slots"])


def test_golden_closest_hits_bit_exact(gpu, golden_small):
    g = golden_small
    log("golden closest: upload reference tree, 4096 rays")
    tree, otree = golden_upload(gpu, g)
    hits = run_closest(gpu, tree, g["rays"])
    want = O.trace(otree, g["positions"], g["indices"], g["rays"])
    assert np.array_equal(hits, want)
    # against the reference's own (native rcpss) hits: same ids but for rounding boundaries
    assert (hits[:, 3] != g["hits"][:, 3]).sum() <= len(hits) // 1000


def test_golden_any_hit_exact(gpu, golden_small):
    g = golden_small
    log("golden any-hit")
    tree, otree = golden_upload(gpu, g)
    occ = run_any(gpu, tree, g["rays"])
    want = O.trace(otree, g["positions"], g["indices"], g["rays"], any_hit=True)
    assert np.array_equal(occ, want)
    assert np.array_equal(occ, g["occluded"])


@pytest.mark.parametrize("any_hit", [False, True])
def test_reference_loop_counts_match_oracle(gpu, golden_small, any_hit):
    import torch

    g = golden_small
    log(f"counts any_hit={any_hit}")
    tree, otree = golden_upload(gpu, g)
    out, cnt = tree.counts(to_dev(g["rays"]), any_hit=any_hit)
    torch.cuda.synchronize()
    want, wcnt = O.trace(otree, g["positions"], g["indices"], g["rays"], any_hit=any_hit,
                         counts=True)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want)
    assert np.array_equal(cnt.cpu().numpy().view(np.uint32), wcnt)


@pytest.fixture(scope="module")
def medium(gpu):
    from sprout_amd import scenes

    pos, idx = scenes.displaced_icosphere(subdivisions=5, num_random=20000, log_mu=-4.0)
    otree = O.build(pos, idx)
    cam = scenes.camera_rays(128, 128)
    chit = O.trace(otree, pos, idx, cam)
    bounce = scenes.bounce_rays(cam, chit, pos, idx, 32768, seed=11)
    rng = np.random.default_rng(5)
    n = 16384
    o = (rng.random((n, 3)) * 2.4 - 1.2).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    rand = np.concatenate([o, np.zeros((n, 1), np.float32), d, np.full((n, 1), 3.4e38, np.float32)], 1)
    rays = np.concatenate([cam, bounce, rand]).astype(np.float32)
    tree = gpu.upload(otree.nodes, otree.slots, pos, idx)
    return pos, idx, otree, tree, rays


def test_medium_mesh_closest_and_any(gpu, medium):
    pos, idx, otree, tree, rays = medium
    log(f"medium: {len(idx)} triangles, {len(rays)} rays")
    assert np.array_equal(run_closest(gpu, tree, rays), O.trace(otree, pos, idx, rays))
    assert np.array_equal(run_any(gpu, tree, rays), O.trace(otree, pos, idx, rays, any_hit=True))


@pytest.mark.parametrize("kernel,block,per_cu,lanes,batch,lds", [
    (0, 64, 1, 1, 1, 32), (0, 128, 3, 8, 7, 4), (0, 256, 8, 16, 256, 16), (0, 512, 2, 64, 64, 1),
    (0, 1024, 1, 32, 1000, 8), (1, 64, 1, 0, 0, 32), (1, 256, 8, 0, 0, 2), (1, 1024, 1, 0, 0, 16),
    (2, 64, 2, 24, 64, 3), (2, 256, 8, 24, 64, 16), (2, 512, 4, 1, 5, 32),
    (3, 64, 2, 24, 64, 3), (3, 256, 8, 24, 64, 16), (3, 256, 6, 24, 64, 10), (3, 512, 4, 1, 5, 32),
    (3, 1024, 1, 32, 1000, 32), (4, 64, 2, 24, 64, 3), (4, 256, 8, 8, 64, 10),
    (4, 512, 4, 1, 5, 32)])
def test_launch_shapes_do_not_change_results(gpu, medium, kernel, block, per_cu, lanes, batch, lds):
    """Every kernel variant and launch shape gives the same bits; small LDS
    stacks force the global spill path on this depth-20 tree."""
    pos, idx, otree, tree, rays = medium
    log(f"kernel {kernel} block={block} blocks_per_cu={per_cu} refill_lanes={lanes} "
        f"claim_batch={batch} lds_entries={lds}")
    gpu.set_trace_config(block, per_cu, kernel, lanes, batch, lds)
    try:
        assert np.array_equal(run_closest(gpu, tree, rays), O.trace(otree, pos, idx, rays))
        assert np.array_equal(run_any(gpu, tree, rays), O.trace(otree, pos, idx, rays, any_hit=True))
    finally:
        gpu.set_trace_config()


@pytest.mark.parametrize("kernel", [3, 4])
def test_sorted_rays_give_the_same_results(gpu, medium, golden_small, kernel):
    """sort_rays: the rays are traced in the order of their octant and
    origin Morton code, their results written in the callers' order: the
    same bits as the oracle, also for one ray and for two."""
    pos, idx, otree, tree, rays = medium
    log(f"kernel {kernel}, rays sorted")
    gpu.set_trace_config(kernel=kernel, sort_rays=1)
    try:
        assert np.array_equal(run_closest(gpu, tree, rays), O.trace(otree, pos, idx, rays))
        assert np.array_equal(run_any(gpu, tree, rays), O.trace(otree, pos, idx, rays, any_hit=True))
        for n in (1, 2, 3):
            assert np.array_equal(run_closest(gpu, tree, rays[-n:]),
                                  O.trace(otree, pos, idx, rays[-n:]))
    finally:
        gpu.set_trace_config()


def test_empty_and_single_ray_batches(gpu, golden_small):
    import torch

    g = golden_small
    log("empty batch")
    tree, otree = golden_upload(gpu, g)
    empty = torch.empty((0, 8), dtype=torch.float32, device="cuda")
    assert tree.intersect(empty).shape == (0, 4)
    assert tree.intersect_p(empty).shape == (0,)
    one = g["rays"][:1]
    assert np.array_equal(run_closest(gpu, tree, one),
                          O.trace(otree, g["positions"], g["indices"], one))


def deep_mesh():
    x = (1.1 ** np.arange(900)).astype(np.float32)
    tris = np.stack([np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1),
                     np.stack([x, np.ones_like(x), np.zeros_like(x)], 1),
                     np.stack([x, np.zeros_like(x), np.ones_like(x)], 1)], 1)
    pos = tris.reshape(-1, 3).astype(np.float32)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    return pos, idx


def test_deep_tree_uses_the_spill_stack(gpu):
    pos, idx = deep_mesh()
    otree = O.build(pos, idx)
    assert otree.depth() > 34  # deeper than the 32 LDS entries
    log(f"deep tree: depth {otree.depth()}")
    tree = gpu.upload(otree.nodes, otree.slots, pos, idx)
    rng = np.random.default_rng(9)
    n = 8192
    o = np.stack([rng.uniform(-1, 1e4, n), rng.uniform(-0.5, 1.5, n), rng.uniform(-0.5, 1.5, n)], 1)
    d = rng.normal(size=(n, 3))
    d[: n // 2, 0] = np.abs(d[: n // 2, 0]) * 20  # mostly along the chain
    rays = np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 3.4e38)], 1).astype(np.float32)
    assert np.array_equal(run_closest(gpu, tree, rays), O.trace(otree, pos, idx, rays))
    assert np.array_equal(run_any(gpu, tree, rays), O.trace(otree, pos, idx, rays, any_hit=True))


def chain_tree(depth):
    """A valid caterpillar tree: inner node 2k has children 2k+1 (leaf) and 2k+2."""
    n = 2 * depth + 1
    nodes = np.zeros(n, O.NODE_DTYPE)
    nodes["lo"] = -1.0
    nodes["hi"] = 1.0
    for k in range(depth):
        nodes["data"][2 * k] = 2 * k + 1
        nodes["axis"][2 * k] = k % 3
        nodes["num"][2 * k] = 0
        nodes["data"][2 * k + 1] = 0
        nodes["num"][2 * k + 1] = 1
    nodes["data"][n - 1] = 0
    nodes["num"][n - 1] = 1
    return nodes


def test_tree_deeper_than_the_reference_stack_is_rejected(gpu):
    import torch

    from sprout_amd._lib import MiError

    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    slots = np.zeros(1, np.uint32)
    ok = gpu.upload(chain_tree(100), slots, pos, idx)
    assert ok.info().max_depth == 100
    rays = torch.tensor([[0.2, 0.2, -1, 0, 0, 0, 1, 3.4e38]], dtype=torch.float32, device="cuda")
    hits = ok.intersect(rays)
    torch.cuda.synchronize()
    assert hits.cpu().numpy().view(np.uint32)[0, 3] == 0
    deep = gpu.upload(chain_tree(126), slots, pos, idx)
    with pytest.raises(MiError) as e:
        deep.intersect(rays)
    assert e.value.status == -6


def test_upload_rejects_malformed_trees(gpu):
    from sprout_amd._lib import MiError

    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    nodes = chain_tree(3)
    bad = nodes.copy()
    bad["data"][2] = 1  # child pointing backwards
    with pytest.raises(MiError):
        gpu.upload(bad, np.zeros(1, np.uint32), pos, idx)
    bad = nodes.copy()
    bad["data"][1] = 5  # leaf slots out of range
    with pytest.raises(MiError):
        gpu.upload(bad, np.zeros(1, np.uint32), pos, idx)
    with pytest.raises(MiError):
        gpu.upload(nodes, np.array([7], np.uint32), pos, idx)  # triangle out of range


def test_autotune_picks_a_shape_that_keeps_the_results(gpu, medium):
    from sprout_amd import bvh

    pos, idx, otree, tree, rays = medium
    log("autotune on the medium mesh")
    try:
        shape, ms = bvh.autotune(tree, rays, 2)
        assert ms > 0
        assert shape["kernel"] == 3 and shape["block_size"] == 256
        assert 1 <= shape["lds_entries"] <= 16 and 4 <= shape["refill_lanes"] <= 24
        assert np.array_equal(run_closest(gpu, tree, rays), O.trace(otree, pos, idx, rays))
        assert np.array_equal(run_any(gpu, tree, rays), O.trace(otree, pos, idx, rays, any_hit=True))
    finally:
        gpu.set_trace_config()


def test_compressed_pairs_on_the_full_mesh(gpu):
    """Kernel 4 (32-byte compressed sibling pairs) against kernel 3 on the
    micro-benchmark's 1.41 M-triangle mesh, 2^20 camera rays and their
    cosine bounce rays: the same hits bit for bit, closest and any."""
    import torch

    from sprout_amd import scenes

    log("displaced icosphere, GPU SBVH")
    pos, idx = scenes.displaced_icosphere()
    tree = gpu.Builder_SAH(16, 64, 4).build(torch.from_numpy(pos).cuda(),
                                            torch.from_numpy(idx.view(np.int32)).cuda())
    cam = scenes.camera_rays(1024, 1024)
    cam_hits = run_closest(gpu, tree, cam)
    bounce = scenes.bounce_rays(cam, cam_hits.view(np.float32).reshape(-1, 4), pos, idx,
                                1024 * 1024, seed=3)
    out = {}
    try:
        for kernel in (3, 4):
            gpu.set_trace_config(kernel=kernel)
            log(f"kernel {kernel}: camera and bounce rays")
            out[kernel] = [run_closest(gpu, tree, cam), run_closest(gpu, tree, bounce),
                           run_any(gpu, tree, bounce)]
    finally:
        gpu.set_trace_config()
    for a, b in zip(out[3], out[4]):
        mism = int(np.count_nonzero(np.any(a.reshape(len(a), -1) != b.reshape(len(b), -1), axis=1)))
        assert 0 == mism, mism
    assert (out[3][1].reshape(-1, 4)[:, 3] != 0xFFFFFFFF).mean() > 0.1  # 22 % of bounces hit
