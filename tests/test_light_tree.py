"""Light tree: the host builder and the oracle's sampling restatement (CPU).

The builder (mi_light_tree_build) is checked structurally (every light once,
orders inverse to the mapping, leaves of at most 4 lights, node power the sum
of its children's, bounds enclosing the children, `middle` the end of the
left subtree) and for the choice of every split: a numpy restatement of the
reference's candidate set and cost (light_tree_builder.cpp:74-83, 149-234,
346-420) must find no cheaper candidate.  The sampling restatement
(oracle/light_tree.c) is held to what the mathematics and the reference's own
debug check require: pdf(pick) equals Tree::pdf of the picked light
(scene.cpp:158-168), pdfs over all lights sum to one, pick frequencies follow
the pdf (chi-square), and with splitting the estimator sum 1/pdf over the
picks is unbiased for the number of lights.
"""

import math

import numpy as np
import pytest

from oracle import oracle as O
from sprout_amd import lights

F = np.float32


def random_lights(n, seed, power_scale=1.0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-2, 2, (n, 3)).astype(F)
    size = rng.uniform(0.01, 0.2, (n, 3)).astype(F)
    lo = (centers - size).astype(F)
    hi = (centers + size).astype(F)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cones = np.concatenate([nrm, np.ones((n, 1))], 1).astype(F)
    powers = (power_scale * np.exp(rng.normal(size=n))).astype(F)
    return lo, hi, cones, powers, np.zeros(n, np.uint8)


def records(lo, hi, cones, powers):
    rec = np.zeros(len(powers), O.LIGHT_DTYPE)
    rec["center"] = F(0.5) * (lo + hi)
    ext = (hi - lo).astype(F)
    rec["radius"] = F(0.5) * np.sqrt((ext[:, 0] * ext[:, 0] + ext[:, 1] * ext[:, 1]) + ext[:, 2] * ext[:, 2])
    rec["cone"] = cones
    rec["power"] = powers
    rec["triangle"] = np.arange(len(powers))
    rec["area"] = 1.0
    return rec


# ---- numpy restatement of the split cost ----

def cone_cost(c):
    o = math.acos(max(-1.0, min(1.0, c)))
    w = min(o + math.pi / 2, math.pi)
    s = math.sin(o)
    b = (math.pi / 2) * (2 * w * s - math.cos(o - 2 * w) - 2 * o * s + c)
    return 2 * math.pi * (1 - c) + b


def cone_merge(a, b):
    if a is None:
        return b
    if np.array_equal(a, b):
        return a
    aa, ba = math.acos(a[3]), math.acos(b[3])
    if ba > aa:
        a, b, aa, ba = b, a, ba, aa
    d = math.acos(max(-1.0, min(1.0, float(np.dot(a[:3], b[:3])))))
    if min(d + ba, math.pi) <= aa:
        return a
    o = (aa + d + ba) / 2
    r = o - aa
    c = np.cross(a[:3], b[:3])
    if not np.dot(c, c) > 0:
        return np.array([*a[:3], -1.0])  # opposite axes: the whole sphere
    c /= np.linalg.norm(c)
    # set_rotation(m, c, r) then transform_vector(m, a): the row vector a
    # times the matrix (base/math/matrix3x3.inl:49-53, 106-132)
    cs, sn, t = math.cos(r), math.sin(r), 1 - math.cos(r)
    m = np.empty((3, 3))
    m[0, 0] = cs + c[0] * c[0] * t
    m[1, 1] = cs + c[1] * c[1] * t
    m[2, 2] = cs + c[2] * c[2] * t
    m[1, 0] = c[0] * c[1] * t + c[2] * sn
    m[0, 1] = c[0] * c[1] * t - c[2] * sn
    m[2, 0] = c[0] * c[2] * t - c[1] * sn
    m[0, 2] = c[0] * c[2] * t + c[1] * sn
    m[2, 1] = c[1] * c[2] * t + c[0] * sn
    m[1, 2] = c[1] * c[2] * t - c[0] * sn
    rot = a[:3] @ m
    rot /= np.linalg.norm(rot)
    return np.array([*rot, math.cos(o)])


def candidates(idx, lo, hi):
    blo = lo[idx].min(0).astype(np.float64)
    bhi = hi[idx].max(0).astype(np.float64)
    ext = bhi - blo
    cands = []
    if len(idx) < 128:
        for i in idx:
            for a in range(3):
                cands.append((a, float(hi[i][a])))
    else:
        la = int(np.argmax(ext))
        step = ext[la] / 16
        for a in range(3):
            ns = int(math.ceil(ext[a] / step))
            for k in range(1, ns):
                cands.append((a, blo[a] + k * (ext[a] / ns)))
    return cands


def split_cost(idx, a, d, lo, hi, cones, powers):
    """Split_candidate::evaluate (scene variant), cones merged in idx order."""
    blo = lo[idx].min(0).astype(np.float64)
    bhi = hi[idx].max(0).astype(np.float64)
    ext = bhi - blo
    sa = 2 * (ext[0] * ext[1] + ext[0] * ext[2] + ext[1] * ext[2])
    cone = None
    for i in idx:
        cone = cone_merge(cone, cones[i].astype(np.float64))
    side0 = [i for i in idx if hi[i][a] < d]
    side1 = [i for i in idx if not hi[i][a] < d]
    reg = ext.max() / ext[a] if ext[a] > 0 else math.inf
    if not side0 or not side1:
        return math.inf
    parts = []
    for side in (side0, side1):
        c = None
        for i in side:
            c = cone_merge(c, cones[i].astype(np.float64))
        e = hi[side].max(0).astype(np.float64) - lo[side].min(0).astype(np.float64)
        s = 2 * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2])
        parts.append(float(powers[side].astype(np.float64).sum()) * cone_cost(c[3]) * s)
    return reg * (parts[0] + parts[1]) / (sa * cone_cost(cone[3]))


def std_partition(seq, pred):
    """libstdc++'s std::partition on bidirectional iterators (swaps from both ends)."""
    seq = list(seq)
    first, last = 0, len(seq)
    while True:
        while True:
            if first == last:
                return seq, first
            if pred(seq[first]):
                first += 1
            else:
                break
        last -= 1
        while True:
            if first == last:
                return seq, first
            if not pred(seq[last]):
                last -= 1
            else:
                break
        seq[first], seq[last] = seq[last], seq[first]
        first += 1


def check_tree(nodes, middles, mapping, orders, lo, hi, cones, powers, check_costs):
    n = len(powers)
    assert sorted(mapping.tolist()) == list(range(n))
    assert np.array_equal(orders[mapping], np.arange(n))

    def bound(nid, begin):
        nd = nodes[nid]
        idx = mapping[begin:begin + nd["num_lights"]]
        blo, bhi = lo[idx].min(0), hi[idx].max(0)
        # the node's sphere is the box of its lights: center, half diagonal
        assert np.allclose(nd["center"], 0.5 * (blo + bhi), rtol=1e-6, atol=1e-6)
        assert np.isclose(nd["radius"], 0.5 * np.linalg.norm(bhi - blo), rtol=1e-5)
        assert np.isclose(nd["power"], powers[idx].astype(np.float64).sum(), rtol=1e-5)

    def walk(nid, begin):
        nd = nodes[nid]
        has = nd["bits"] & 1
        child = nd["bits"] >> 2
        bound(nid, begin)
        if not has:
            assert child == begin and 1 <= nd["num_lights"] <= 4
            assert middles[nid] == 0
            return begin + nd["num_lights"]
        c0, c1 = nodes[child], nodes[child + 1]
        assert nd["num_lights"] == c0["num_lights"] + c1["num_lights"]
        mid = walk(child, begin)
        assert middles[nid] == mid
        end = walk(child + 1, mid)
        return end

    assert walk(0, 0) == n
    if check_costs:
        replay(nodes, middles, mapping, lo, hi, cones, powers)


def replay(nodes, middles, mapping, lo, hi, cones, powers):
    """Rebuilds the light order the reference's recursion produces: at every
    inner node the split the builder took must be a cheapest candidate
    (candidate costs recomputed here with cones merged in the current order),
    and partitioning with it must give the builder's order."""
    order = list(range(len(powers)))

    def rec(nid, begin, end):
        nd = nodes[nid]
        if not (nd["bits"] & 1):
            return
        idx = order[begin:end]
        mid = int(middles[nid])
        left = set(mapping[begin:mid].tolist())
        costs = [(split_cost(idx, a, d, lo, hi, cones, powers), a, d)
                 for a, d in candidates(idx, lo, hi)]
        best = min(c for c, _, _ in costs)
        taken = [(c, a, d) for c, a, d in costs
                 if set(i for i in idx if hi[i][a] < d) == left]
        assert taken, "the builder's split is none of the candidates"
        c, a, d = min(taken)
        assert c <= best * (1 + 1e-4), (c, best)
        part, k = std_partition(idx, lambda i: hi[i][a] < d)
        assert begin + k == mid
        order[begin:end] = part
        child = int(nd["bits"] >> 2)
        rec(child, begin, mid)
        rec(child + 1, mid, end)

    rec(0, 0, len(order))
    assert order == mapping.tolist()


@pytest.mark.parametrize("n,seed,costs", [(1, 0, False), (2, 1, True), (5, 2, True),
                                          (40, 3, True), (300, 4, False)])
def test_builder_structure_and_splits(n, seed, costs):
    lo, hi, cones, powers, ts = random_lights(n, seed)
    nodes, middles, mapping, orders = lights.build_tree(lo, hi, cones, powers, ts)
    assert len(nodes) <= 2 * n - 1
    check_tree(nodes, middles, mapping, orders, lo, hi, cones, powers, costs)


def test_builder_on_grid_candidates_is_consistent():
    """300 lights take the 16-slice grid path; the structure checks apply."""
    lo, hi, cones, powers, ts = random_lights(300, 9)
    nodes, middles, mapping, orders = lights.build_tree(lo, hi, cones, powers, ts)
    check_tree(nodes, middles, mapping, orders, lo, hi, cones, powers, False)
    # the root split uses a grid candidate: one of its planes at a slice position
    assert nodes[0]["bits"] & 1


def make_tree(n, seed, power_scale=1.0):
    lo, hi, cones, powers, ts = random_lights(n, seed, power_scale)
    nodes, middles, mapping, orders = lights.build_tree(lo, hi, cones, powers, ts)
    return O.LightTree(nodes, middles, mapping, orders, records(lo, hi, cones, powers))


def shading_points(m, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-3, 3, (m, 3)).astype(F)
    nrm = rng.normal(size=(m, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return p, nrm.astype(F), rng


@pytest.mark.parametrize("n", [1, 3, 17, 120])
def test_pdfs_sum_to_one_and_match_picks(n):
    tree = make_tree(n, n)
    p, nrm, rng = shading_points(20, 5)
    for k in range(len(p)):
        pdfs = np.array([tree.pdf(p[k], nrm[k], False, l) for l in range(n)])
        assert abs(pdfs.sum() - 1) < 1e-5, pdfs.sum()
        for _ in range(20):
            ls, ps = tree.random_light(p[k], nrm[k], float(rng.random()), False)
            assert len(ls) == 1
            assert abs(ps[0] - pdfs[ls[0]]) <= 1e-6 * pdfs[ls[0]] + 1e-12


def test_pick_frequencies_follow_pdf():
    n = 24
    tree = make_tree(n, 11)
    p = np.array([0.3, -0.2, 0.1], F)
    nrm = np.array([0, 0, 1], F)
    pdfs = np.array([tree.pdf(p, nrm, False, l) for l in range(n)])
    m = 40000
    u = (np.arange(m) + 0.5) / m  # stratified uniforms
    counts = np.zeros(n)
    for x in u:
        ls, _ = tree.random_light(p, nrm, float(x), False)
        counts[ls[0]] += 1
    expected = pdfs * m
    big = expected > 20
    chi2 = ((counts[big] - expected[big]) ** 2 / expected[big]).sum()
    # stratified draws: far below the chi-square bound of independent ones
    assert chi2 < 3 * big.sum(), chi2


def test_split_picks_follow_their_pdf():
    """With splitting several lights come back per query; each light is
    returned with the probability its pdf states (the marginal the MIS
    weights need), and that pdf is Tree::pdf with splitting."""
    n = 60
    # bright lights and points close to them make the variance test split nodes
    tree = make_tree(n, 13, power_scale=100.0)
    p, nrm, _ = shading_points(6, 17)
    near = tree.lights["center"][:4] + np.float32(0.01)
    p = np.concatenate([p, near]).astype(F)
    nrm = np.concatenate([nrm, np.tile(np.array([[0, 0, 1]], F), (4, 1))]).astype(F)
    m = 20000
    total_multi = 0
    for k in range(len(p)):
        freq = np.zeros(n)
        multi = 0
        for j in range(m):
            ls, ps = tree.random_light(p[k], nrm[k], (j + 0.5) / m, True)
            assert len(set(ls.tolist())) == len(ls)
            multi += len(ls) > 1
            for l, q in zip(ls, ps):
                freq[l] += 1
                assert abs(q - tree.pdf(p[k], nrm[k], True, l)) <= 1e-6 * q
        freq /= m
        pdfs = np.array([tree.pdf(p[k], nrm[k], True, l) for l in range(n)])
        tol = 5 * np.sqrt(pdfs * (1 - pdfs) / m) + 2.0 / m
        assert np.all(np.abs(freq - pdfs) <= tol), (k, np.abs(freq - pdfs).max())
        total_multi += multi
    assert total_multi > 0, "no query split the tree"


def test_opposite_and_identical_lights_build():
    """Lights facing opposite ways in one node (the reference's cone merge
    divides by zero there) and lights with identical boxes (every candidate
    leaves a side empty) still give a valid tree whose pdfs sum to one."""
    n = 24
    lo = np.zeros((n, 3), F)
    hi = np.full((n, 3), 0.1, F)
    lo[: n // 2] += F(0.5)  # half the lights share one box, half another ... and
    hi[: n // 2] += F(0.5)
    lo[:6] = lo[6]  # six more identical ones
    hi[:6] = hi[6]
    cones = np.zeros((n, 4), F)
    cones[:, 1] = np.where(np.arange(n) % 2 == 0, 1, -1)
    cones[:, 3] = 1
    powers = np.ones(n, F)
    nodes, middles, mapping, orders = lights.build_tree(lo, hi, cones, powers, np.zeros(n, np.uint8))
    check_tree(nodes, middles, mapping, orders, lo, hi, cones, powers, False)
    assert np.all(np.isfinite(nodes["cone"]))
    tree = O.LightTree(nodes, middles, mapping, orders, records(lo, hi, cones, powers))
    p = np.array([0.3, 1.0, 0.2], F)
    nrm = np.array([0, -1, 0], F)
    assert abs(sum(tree.pdf(p, nrm, False, l) for l in range(n)) - 1) < 1e-5

