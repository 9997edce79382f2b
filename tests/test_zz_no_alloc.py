"""mi_trace_closest_noalloc: closest hits without any device allocation.

The entry point takes its ray counter and traversal-stack spill area from the
library's fixed device array instead of allocating them, and shrinks the
persistent grid to what that array holds.  Through the C ABI it must return
exactly what mi_trace_closest returns, on two calls in a row, the second with
a smaller input (the array keeps the first call's contents).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def lib():
    import torch

    from sprout_amd import _lib, bvh

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield _lib.load()
    torch.cuda.synchronize()


def chain_mesh():
    """900 unit triangles along x at geometric spacing: an SBVH deeper than
    the 10 stack entries a lane keeps in LDS, so traversals spill."""
    x = (1.1 ** np.arange(900)).astype(np.float32)
    tris = np.stack([np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1),
                     np.stack([x, np.ones_like(x), np.zeros_like(x)], 1),
                     np.stack([x, np.zeros_like(x), np.ones_like(x)], 1)], 1)
    pos = tris.reshape(-1, 3).astype(np.float32)
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    return pos, idx


def chain_rays(n, seed):
    """Half random rays, half aimed along +x at the inside of one of the first
    200 triangles (x below 2e8, where float32 keeps the aim)."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-1, 1e4, n), rng.uniform(-0.5, 1.5, n), rng.uniform(-0.5, 1.5, n)], 1)
    d = rng.normal(size=(n, 3))
    h = n // 2
    x = 1.1 ** rng.integers(0, 200, h)
    o[:h] = np.stack([x * 0.5 - 1.0, rng.uniform(0.05, 0.45, h), rng.uniform(0.05, 0.45, h)], 1)
    d[:h] = np.stack([np.ones(h), rng.normal(0, 1e-3, h), rng.normal(0, 1e-3, h)], 1)
    return np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 3.4e38)], 1).astype(np.float32)


def trace_both(lib, handle, rays):
    """(mi_trace_closest hits, mi_trace_closest_noalloc hits) as uint32 words."""
    import torch

    from sprout_amd._lib import check

    d = torch.from_numpy(rays).cuda()
    n = d.shape[0]
    want = torch.full((n, 4), -1, dtype=torch.int32, device="cuda")
    got = torch.full((n, 4), -7, dtype=torch.int32, device="cuda")
    check(lib.mi_trace_closest(handle, d.data_ptr(), n, want.data_ptr(), None), "mi_trace_closest")
    check(lib.mi_trace_closest_noalloc(handle, d.data_ptr(), n, got.data_ptr(), None),
          "mi_trace_closest_noalloc")
    torch.cuda.synchronize()
    return want.cpu().numpy().view(np.uint32), got.cpu().numpy().view(np.uint32)


def test_noalloc_equals_closest_on_two_calls(lib):
    """A spilling tree and 2^19 rays: the usual grid (8 workgroups per CU)
    spills more than the fixed array holds, so the call runs on a smaller
    persistent grid; then 1 000 rays reuse the array.  Both equal
    mi_trace_closest bit for bit."""
    import torch

    from sprout_amd import bvh

    pos, idx = chain_mesh()
    tree = bvh.Builder_SAH().build(torch.from_numpy(pos).cuda(),
                                   torch.from_numpy(idx.view(np.int32)).cuda())
    info = tree.info()
    log(f"chain tree: {info.num_nodes} nodes, depth {info.max_depth}")
    assert info.max_depth > 20
    for n, seed in ((1 << 19, 3), (1000, 4)):
        log(f"  {n} rays")
        want, got = trace_both(lib, tree.handle, chain_rays(n, seed))
        bad = np.nonzero(np.any(want != got, axis=1))[0]
        hit = float(np.mean(want[:, 3] != 0xFFFFFFFF))
        log(f"    hit fraction {hit:.3f}")
        assert len(bad) == 0, f"{len(bad)} of {n} rays differ, first {bad[:4]}"
        assert hit > 0.15


def test_noalloc_empty_batch_and_bad_handle(lib):
    from sprout_amd._lib import MI_NULL

    assert lib.mi_trace_closest_noalloc(0, None, 0, None, None) in (0, -4)
    assert lib.mi_trace_closest_noalloc(MI_NULL - 1, None, 1, None, None) == -4
