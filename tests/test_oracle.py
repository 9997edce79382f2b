"""Pins the CPU oracle (oracle/) before it is trusted as the GPU checker.

The oracle restates sprout's SBVH builder and traversal.  It is checked
against (a) the golden fixtures written from the reference's own code
(tests/golden/make_golden.py), (b) the reference library itself when it was
built here (oracle/_ref), (c) brute force over all triangles and (d) closed
forms of the arithmetic it pins.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O

NULL = 0xFFFFFFFF


def golden_tree(g):
    nodes = np.ascontiguousarray(g["nodes"]).view(O.NODE_DTYPE).reshape(-1)
    return O.Tree(nodes, g["slots"])


def native_rcp_matches_host(g) -> bool:
    here = np.array([O.lib().oracle_reciprocal(float(x), O.RCP_NATIVE) for x in g["rcp_probe"]],
                    np.float32)
    return bool(np.array_equal(here.view(np.uint32), g["rcp_native"].view(np.uint32)))


def test_oracle_builds_the_reference_tree(golden_small):
    g = golden_small
    tree = O.build(g["positions"], g["indices"])
    ref = golden_tree(g)
    assert tree.num_nodes == ref.num_nodes
    assert O.nodes_equal(tree.nodes, ref.nodes)
    assert np.array_equal(tree.slots, ref.slots)


def test_native_traversal_matches_reference_bits(golden_small):
    g = golden_small
    if not native_rcp_matches_host(g):
        pytest.skip("this host's rcpss differs from the host that wrote the fixture")
    tree = golden_tree(g)
    hits = O.trace(tree, g["positions"], g["indices"], g["rays"], rcp_mode=O.RCP_NATIVE)
    assert np.array_equal(hits, g["hits"])
    occ = O.trace(tree, g["positions"], g["indices"], g["rays"], any_hit=True,
                  rcp_mode=O.RCP_NATIVE)
    assert np.array_equal(occ, g["occluded"])


def test_pinned_reciprocal_stays_within_tolerance_of_reference(golden_small):
    """Gate: oracle(pinned) vs the reference (native rcpss): same hit ids on all
    but rounding-boundary rays, t within 1e-6 relative where ids agree."""
    g = golden_small
    tree = golden_tree(g)
    hits = O.trace(tree, g["positions"], g["indices"], g["rays"], rcp_mode=O.RCP_PINNED)
    ref = g["hits"]
    same = hits[:, 3] == ref[:, 3]
    assert (~same).sum() <= max(1, len(ref) // 1000)
    hit = same & (ref[:, 3] != NULL)
    t = hits[hit, 0].view(np.float32)
    tr = ref[hit, 0].view(np.float32)
    assert np.all(np.abs(t - tr) <= 1e-6 * np.abs(tr))


def test_traversal_agrees_with_brute_force(golden_small):
    """Closest hit is independent of the tree: the same primitive test over all
    triangles in input order must give the same t; ids differ only on ties."""
    g = golden_small
    tree = golden_tree(g)
    hits = O.trace(tree, g["positions"], g["indices"], g["rays"])
    bf = O.brute_force(g["positions"], g["indices"], g["rays"])
    assert np.array_equal(hits[:, 0], bf[:, 0])
    diff = hits[:, 3] != bf[:, 3]
    # a different id with bit-identical t is a tie between triangles
    assert np.all(hits[diff, 0] == bf[diff, 0])
    assert diff.sum() <= len(hits) // 100
    miss = bf[:, 3] == NULL
    assert np.array_equal(hits[:, 3] == NULL, miss)


def test_any_hit_is_consistent_with_closest_hit(golden_small):
    g = golden_small
    tree = golden_tree(g)
    hits = O.trace(tree, g["positions"], g["indices"], g["rays"])
    occ = O.trace(tree, g["positions"], g["indices"], g["rays"], any_hit=True)
    assert np.array_equal(occ.astype(bool), hits[:, 3] != NULL)


def test_tree_structure_invariants(golden_small):
    g = golden_small
    tree = O.build(g["positions"], g["indices"])
    nodes = tree.nodes
    inner = nodes["num"] == 0
    assert len(nodes) == 2 * inner.sum() + 1
    # every triangle is referenced by a leaf
    assert np.array_equal(np.unique(tree.slots), np.arange(len(g["indices"])))
    # leaves cover the slot array exactly once, in order
    leaves = np.nonzero(~inner)[0]
    starts = nodes["data"][leaves]
    order = np.argsort(starts)
    ends = starts[order] + nodes["num"][leaves][order]
    assert starts[order][0] == 0 and np.array_equal(starts[order][1:], ends[:-1])
    assert ends[-1] == len(tree.slots)
    # children lie inside their parent's box
    for i in np.nonzero(inner)[0]:
        c = nodes["data"][i]
        assert c > i
        for k in (c, c + 1):
            assert np.all(nodes["lo"][k] >= nodes["lo"][i]) and np.all(nodes["hi"][k] <= nodes["hi"][i])
    # leaf triangles lie inside (or, clipped by spatial splits, overlap) their leaf box
    pos = g["positions"]
    idx = g["indices"]
    for li in leaves[:200]:
        for s in range(nodes["data"][li], nodes["data"][li] + nodes["num"][li]):
            tri = pos[idx[tree.slots[s]]]
            assert np.all(tri.max(0) >= nodes["lo"][li]) and np.all(tri.min(0) <= nodes["hi"][li])


def test_reciprocal_closed_forms():
    for k in range(-20, 21):
        x = np.float32(2.0 ** k)
        assert O.reciprocal(x) == np.float32(2.0 ** -k)
        assert O.reciprocal3([x, -x, x])[1] == np.float32(-(2.0 ** -k))
    # zero direction lanes are NaN in the reference (inf * 0 in the Newton step)
    assert np.isnan(O.reciprocal3([0.0, 1.0, 1.0])[0])
    assert np.isnan(O.reciprocal3([-0.0, 1.0, 1.0])[0])
    assert np.isnan(O.reciprocal(0.0))
    # r * r overflows for |x| < ~5.4e-20: the fused step gives -inf for a
    # positive lane, so its sign bit (and near/far order) flips
    assert O.reciprocal3([1e-25, 1.0, 1.0])[0] == -np.inf
    xs = np.random.default_rng(1).uniform(0.01, 100, 2000).astype(np.float32)
    r = np.array([O.reciprocal(x) for x in xs], np.float32)
    assert np.all(np.abs(r * xs - 1) < 3e-7)


def test_surface_area_closed_form():
    rng = np.random.default_rng(2)
    for _ in range(200):
        lo = rng.integers(-50, 50, 3).astype(np.float32)
        hi = lo + rng.integers(0, 50, 3).astype(np.float32)
        d = (hi - lo).astype(np.float64)
        assert O.surface_area(lo, hi) == 2 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])


def test_offset_ray_closed_forms():
    one = np.float32(1.0).view(np.uint32)
    out = O.offset_ray([1.0, -1.0, 0.5], [1.0, 1.0, -1.0])
    assert out[0].view(np.uint32) == one + 256
    # p < 0: the integer offset is negated, so the point still moves along +n
    assert out[1].view(np.uint32) == np.float32(-1.0).view(np.uint32) - 256
    assert out[1] > -1.0
    assert out[2].view(np.uint32) == np.float32(0.5).view(np.uint32) - 256
    # near the origin the offset is n / 65536
    out = O.offset_ray([0.0, 0.01, -0.02], [1.0, 0.5, 0.25])
    assert out[0] == np.float32(1.0 / 65536)
    assert out[1] == np.float32(0.01) + np.float32(0.5 / 65536)


REF_LIB = O.ref_lib_path()


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference not built here (oracle/ref.mk)")
@pytest.mark.parametrize("nt,seed", [(40, 1), (700, 2), (5000, 3), (30000, 4)])
def test_oracle_matches_reference_library(nt, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((nt, 3)).astype(np.float32)
    e = (np.exp(rng.normal(-3, 0.7, (nt, 3, 3))) * rng.choice([-1, 1], (nt, 3, 3))).astype(np.float32)
    pos = (c[:, None, :] + e).reshape(-1, 3).astype(np.float32)
    idx = np.arange(nt * 3, dtype=np.uint32).reshape(-1, 3)
    ref = O.RefTree(pos, idx)
    rtree = ref.tree()
    tree = O.build(pos, idx)
    assert O.nodes_equal(tree.nodes, rtree.nodes) and np.array_equal(tree.slots, rtree.slots)

    nr = 8192
    o = (rng.random((nr, 3)) * 1.6 - 0.3).astype(np.float32)
    d = rng.normal(size=(nr, 3)).astype(np.float32)
    rays = np.concatenate([o, np.zeros((nr, 1), np.float32), d,
                           np.full((nr, 1), 3.4e38, np.float32)], 1)
    h_ref, _ = ref.trace(rays)
    a_ref, _ = ref.trace(rays, any_hit=True)
    assert np.array_equal(O.trace(tree, pos, idx, rays, rcp_mode=O.RCP_NATIVE), h_ref)
    assert np.array_equal(O.trace(tree, pos, idx, rays, any_hit=True, rcp_mode=O.RCP_NATIVE), a_ref)
