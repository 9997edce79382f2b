"""The oracle's medium restatement (oracle/medium.c) pinned by what the
mathematics fixes, and the library's host gridtree builder against it:

- density lookups against an independent numpy trilinear / nearest filter;
- the gridtree: the library's builder and the oracle's give identical
  nodes and leaves, and every leaf bounds the filtered density anywhere in
  its box (majorant >= mu >= minorant, checked by brute force);
- ratio / residual ratio tracking: exact exp(-sigma d) in a homogeneous
  medium, and an unbiased mean exp(-integral of sigma) in a layered one;
- RGB delta tracking: in a homogeneous grey medium the free-flight distance
  is exponential (KS test) and the scattering weight is the albedo; in a
  chromatic one the weighted scatter density per channel is
  mu_s,c * exp(-mu_t,c t) and the weighted pass probability exp(-mu_t,c d)
  (the history-aware channel weights are what makes both hold);
- Henyey-Greenstein: normalised, sampled with pdf = phase and mean cosine g;
- the similarity relation's scale over depth;
- the path tracer with a medium: MIS, BSDF-only and NEE-only estimators
  agree (the transmittance of shadow rays and the whole-sphere light-tree
  pdf at medium vertices enter only the first and third).
"""

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O


def lin_density(grid, uvw):
    """numpy trilinear lookup with clamp addressing (texel centres at (i + 0.5) / d)."""
    nz, ny, nx = grid.shape
    dims = np.array([nx, ny, nz], np.float64)
    u = np.clip(np.asarray(uvw, np.float64), 0, 1) * dims - 0.5
    f = np.floor(u)
    i0 = f.astype(int)
    t = u - f
    i1 = np.minimum(i0 + 1, dims.astype(int) - 1)
    i0 = np.maximum(i0, 0)
    val = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2])
                x = i1[0] if dx else i0[0]
                y = i1[1] if dy else i0[1]
                z = i1[2] if dz else i0[2]
                val += w * grid[z, y, x]
    return val


def random_grid(shape, seed, smooth=True):
    rng = np.random.default_rng(seed)
    g = rng.random(shape).astype(np.float32)
    if smooth:
        from scipy.ndimage import gaussian_filter
        g = gaussian_filter(g, 2.0)
        g = (g - g.min()) / (g.max() - g.min())
    g[g < 0.3] = 0.0  # empty regions
    return (g * 3.0).astype(np.float32)


@pytest.mark.parametrize("shape", [(8, 8, 8), (33, 40, 70), (64, 64, 64)])
def test_density_lookups_match_numpy(shape):
    grid = random_grid(shape, 1)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), (1, 1, 1), (1, 1, 1), 0.0)
    rng = np.random.default_rng(2)
    pts = rng.uniform(-0.1, 1.1, (400, 3)).astype(np.float32)
    for p in pts:
        ref = lin_density(grid, p)
        assert abs(m.density_at(p) - ref) <= 2e-6 * max(1.0, abs(ref))
        nz, ny, nx = grid.shape
        q = np.clip(p.astype(np.float64), 0, 1)
        ix = min(int(np.float32(q[0]) * np.float32(nx)), nx - 1)
        iy = min(int(np.float32(q[1]) * np.float32(ny)), ny - 1)
        iz = min(int(np.float32(q[2]) * np.float32(nz)), nz - 1)
        assert m.density_at(p, nearest=True) == grid[iz, iy, ix]


def leaf_boxes(nodes, data, num_cells, dims):
    """(lo voxel, hi voxel, cm) of every non-empty leaf, by walking the tree."""
    out = []

    def walk(index, lo, hi):
        n = int(nodes[index])
        if n & 1:
            first = n >> 1
            c = [lo[k] + ((hi[k] - lo[k]) >> 1) for k in range(3)]
            for i in range(8):
                slo = [c[k] if (i >> k) & 1 else lo[k] for k in range(3)]
                shi = [hi[k] if (i >> k) & 1 else c[k] for k in range(3)]
                walk(first + i, slo, shi)
        elif (n >> 1) != 0x7FFFFFFF:
            out.append((lo, hi, data[n >> 1]))
        else:
            out.append((lo, hi, None))

    for z in range(num_cells[2]):
        for y in range(num_cells[1]):
            for x in range(num_cells[0]):
                i = (z * num_cells[1] + y) * num_cells[0] + x
                lo = [x * 32, y * 32, z * 32]
                walk(i, lo, [v + 32 for v in lo])
    return out


@pytest.mark.parametrize("shape,seed", [((40, 33, 70), 3), ((64, 64, 64), 4), ((17, 5, 100), 5)])
def test_gridtree_library_matches_oracle_and_bounds_density(shape, seed):
    from sprout_amd import medium as M

    grid = random_grid(shape, seed)
    cc_a, cc_s = (0.2, 0.1, 0.3), (0.7, 0.9, 0.5)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
    onodes, odata, ocells = m.gridtree()
    lnodes, ldata, lcells = M.gridtree(grid, cc_a, cc_s)
    assert tuple(ocells) == tuple(int(c) for c in lcells)
    assert np.array_equal(onodes, lnodes)
    assert np.array_equal(odata.view(np.uint32), ldata.view(np.uint32))
    nz, ny, nx = grid.shape
    if min(grid.shape) >= 16:
        assert (onodes & 1).any(), "no cell was refined"
    else:
        # a cell thinner than 9 voxels is never split
        assert not (onodes & 1).any()
    dims = (nx, ny, nz)
    a_lo, a_hi, s_lo, s_hi = min(cc_a), max(cc_a), min(cc_s), max(cc_s)
    leaves = leaf_boxes(onodes, odata, ocells, dims)
    rng = np.random.default_rng(seed)
    for lo, hi, cm in leaves:
        hi_c = [min(hi[k], dims[k]) for k in range(3)]
        if any(hi_c[k] <= lo[k] for k in range(3)):
            continue
        block = grid[lo[2]:hi_c[2], lo[1]:hi_c[1], lo[0]:hi_c[0]]
        if cm is None:
            # an empty leaf: the filtered density is zero in it
            assert block.max() == 0.0 or (a_hi == 0 and s_hi == 0)
            continue
        # random points inside the leaf box: filtered density between the bounds
        for _ in range(8):
            vox = [rng.uniform(lo[k], hi_c[k]) for k in range(3)]
            uvw = np.array([vox[k] / dims[k] for k in range(3)], np.float32)
            d = lin_density(grid, uvw)
            assert d * a_hi <= cm[2] * (1 + 1e-6) + 1e-7
            assert d * s_hi <= cm[3] * (1 + 1e-6) + 1e-7
            assert d * a_lo >= cm[0] * (1 - 1e-6) - 1e-7
            assert d * s_lo >= cm[1] * (1 - 1e-6) - 1e-7


@pytest.mark.parametrize("shape,seed", [((64, 64, 64), 4), ((40, 33, 70), 3), ((17, 5, 100), 5)])
def test_flat_majorants_library_matches_oracle_bound_and_are_tighter(shape, seed):
    """The flat majorant grid (every 32^3-voxel cell split to 8^3-voxel
    leaves): the library's and the oracle's builders agree bit for bit; every
    cell at least 9 voxels wide (with its border) reaches the 8^3 leaves;
    every leaf bounds the filtered density; and at every voxel the flat
    majorant is at most the gridtree's (a leaf's voxels and border lie inside
    its ancestors')."""
    from sprout_amd import medium as M

    grid = random_grid(shape, seed)
    cc_a, cc_s = (0.2, 0.1, 0.3), (0.7, 0.9, 0.5)
    flat = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0, majorants="flat")
    tree = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
    fnodes, fdata, fcells = flat.gridtree()
    lnodes, ldata, lcells = M.gridtree(grid, cc_a, cc_s, majorants="flat")
    assert tuple(fcells) == tuple(int(c) for c in lcells)
    assert np.array_equal(fnodes, lnodes)
    assert np.array_equal(fdata.view(np.uint32), ldata.view(np.uint32))
    with pytest.raises(ValueError):
        M.gridtree(grid, cc_a, cc_s, majorants="octree")

    nz, ny, nx = grid.shape
    dims = (nx, ny, nz)
    fleaves = leaf_boxes(fnodes, fdata, fcells, dims)
    for lo, hi, _ in fleaves:
        span = [min(hi[k], dims[k]) - lo[k] for k in range(3)]
        if min(span) > 0 and all(min(dims[k] - lo[k], 32) >= 9 for k in range(3)):
            assert max(hi[k] - lo[k] for k in range(3)) == 8, (lo, hi)

    def majorant_map(leaves):
        m = np.zeros((nz, ny, nx, 2), np.float64)
        for lo, hi, cm in leaves:
            if cm is not None:
                m[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = cm[2:4]
        return m

    mf = majorant_map(fleaves)
    mt = majorant_map(leaf_boxes(*tree.gridtree()[:2], tree.gridtree()[2], dims))
    assert np.all(mf <= mt)
    # each voxel's own density is under its flat majorant
    for c in range(2):
        hi_cc = max(cc_a) if c == 0 else max(cc_s)
        assert np.all(grid * hi_cc <= mf[..., c] * (1 + 1e-6) + 1e-7)
    if min(shape) >= 16:
        # a low-contrast cloud (density range under the gridtree's 0.1
        # split threshold) keeps the gridtree at its 32^3 cells; the flat
        # grid is strictly tighter there
        faint = grid * np.float32(0.03)
        f = O.Medium(faint, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0, majorants="flat")
        t = O.Medium(faint, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
        assert not (t.gridtree()[0] & 1).any()
        mf = majorant_map(leaf_boxes(*f.gridtree()[:2], f.gridtree()[2], dims))
        mt = majorant_map(leaf_boxes(*t.gridtree()[:2], t.gridtree()[2], dims))
        assert np.all(mf <= mt) and mf.sum() < 0.95 * mt.sum()


@pytest.mark.parametrize("majorants", ["gridtree", "flat"])
def test_ratio_tracking_is_unbiased_in_a_cloud(majorants):
    """A smooth random 64^3 cloud: the mean ratio-tracking estimate along a
    diagonal ray equals exp(-integral of mu_t) (numpy quadrature of the
    trilinear density) with either kind of majorants."""
    grid = random_grid((64, 64, 64), 8)
    cc_a, cc_s = (0.5, 0.8, 0.3), (2.0, 1.5, 2.5)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0, majorants=majorants)
    n = 20000
    o = np.array([-0.5, 0.1, 0.2])
    d = np.array([1.0, 0.7, 0.55])
    d /= np.linalg.norm(d)
    out = m.batch(rays_through(n, o, d), transmittance=True, seed=11)
    # the ray's span inside the unit box
    t_in = max((0 - o[k]) / d[k] for k in range(3))
    t_out = min((1 - o[k]) / d[k] for k in range(3))
    ts = np.linspace(t_in, t_out, 8001)
    mid = 0.5 * (ts[1:] + ts[:-1])
    dens = np.array([lin_density(grid, o + t * d) for t in mid])
    tau = dens.sum() * (ts[1] - ts[0])
    expect = np.exp(-tau * (np.array(cc_a) + np.array(cc_s)))
    mean = out[:, 2:].mean(axis=0)
    sem = out[:, 2:].std(axis=0) / np.sqrt(n)
    assert np.all(np.abs(mean - expect) < 5 * sem + 2e-3 * expect), (mean, expect, sem)


def rays_through(n, o, d, max_t=1e30):
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = max_t
    return r


def test_homogeneous_transmittance():
    """constant density d <= 1: minorant == majorant, residual ratio tracking
    returns exp(-mu_t d) with no random steps.  Above 1 the builder's
    minorant stays at 1 (its running minimum starts at 1,
    volumetric_octree_builder.cpp:176), so the residual is tracked and only
    the mean is exp(-mu_t d)."""
    grid = np.full((16, 16, 16), 0.8, np.float32)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), (0.25, 0.25, 0.25), (0.5, 0.5, 0.5), 0.0)
    nodes, data, _ = m.gridtree()
    assert np.allclose(data[:, 0:2], data[:, 2:4])
    # from below the box straight up through it: length 1, mu_t = 0.8 * 0.75
    r = rays_through(4, (0.5, -1.0, 0.5), (0, 1, 0))
    out = m.batch(r, transmittance=True)
    assert np.all(out[:, 0] == 1)
    t0 = np.float32(O.offset_f(1.0))
    expect = np.exp(-(2.0 - float(t0)) * 0.6)
    assert np.allclose(out[:, 2:], expect, rtol=1e-5)
    # a shadow segment ending inside the box
    r = rays_through(1, (0.5, 0.5, 0.5), (1, 0, 0), 0.25)
    out = m.batch(r, transmittance=True)
    assert np.allclose(out[0, 2:], np.exp(-0.25 * 0.6), rtol=1e-5)

    grid = np.full((16, 16, 16), 2.0, np.float32)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), (0.25, 0.25, 0.25), (0.5, 0.5, 0.5), 0.0)
    nodes, data, _ = m.gridtree()
    assert np.allclose(data[:, 0:2], [0.25, 0.5]) and np.allclose(data[:, 2:4], [0.5, 1.0])
    n = 40000
    out = m.batch(rays_through(n, (0.5, -1.0, 0.5), (0, 1, 0)), transmittance=True, seed=3)
    expect = np.exp(-(2.0 - float(t0)) * 1.5)
    mean = out[:, 2].mean()
    assert abs(mean - expect) < 5 * out[:, 2].std() / np.sqrt(n)


@pytest.mark.parametrize("nearest", [False, True])
def test_ratio_tracking_mean_is_unbiased_in_layered_medium(nearest):
    """Two layers along y (density 0.5 and 4); the mean ratio-tracking
    estimate equals exp(-integral of mu_t) along the ray (integral of the
    filtered density, computed in float64)."""
    ny = 32
    grid = np.zeros((8, ny, 8), np.float32)
    grid[:, : ny // 2, :] = 0.5
    grid[:, ny // 2:, :] = 4.0
    cc_a, cc_s = (0.1, 0.3, 0.2), (0.4, 0.2, 0.6)
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
    n = 20000
    r = rays_through(n, (0.5, -0.5, 0.5), (0, 1, 0))
    out = m.batch(r, transmittance=True, nearest=nearest, seed=7)
    t0 = float(np.float32(O.offset_f(0.5)))
    ys = np.linspace(t0 - 0.5, 1.0, 20001)
    ys_mid = 0.5 * (ys[1:] + ys[:-1])
    if nearest:
        dens = grid[4, np.minimum((ys_mid * ny).astype(int), ny - 1), 4]
    else:
        dens = np.array([lin_density(grid, (0.5, y, 0.5)) for y in ys_mid])
    tau = np.sum(dens) * (ys[1] - ys[0])
    mu = np.array(cc_a) + np.array(cc_s)
    expect = np.exp(-tau * mu)
    mean = out[:, 2:].mean(axis=0)
    sem = out[:, 2:].std(axis=0) / np.sqrt(n)
    assert np.all(np.abs(mean - expect) < 5 * sem + 1e-4), (mean, expect, sem)
    assert (out[:, 2:].std(axis=0) > 0).all(), "the layered medium was not ratio tracked"


def test_delta_tracking_free_flight_is_exponential():
    """grey homogeneous medium: P(scatter before d) = 1 - exp(-mu_t d), the
    distance is exponential (KS test) and the weight is the albedo"""
    grid = np.full((8, 8, 8), 1.0, np.float32)
    mu_a, mu_s = 0.6, 1.4
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), (mu_a,) * 3, (mu_s,) * 3, 0.0)
    n = 20000
    r = rays_through(n, (0.5, 0.5, -1.0), (0, 0, 1))
    out = m.batch(r, seed=3)
    t0 = float(np.float32(O.offset_f(1.0)))
    d = 2.0 - t0
    mu_t = mu_a + mu_s
    scattered = out[:, 0] == 1
    p = 1 - np.exp(-mu_t * d)
    assert abs(scattered.mean() - p) < 5 * np.sqrt(p * (1 - p) / n)
    assert np.allclose(out[scattered, 2:], mu_s / mu_t, rtol=1e-6)
    assert np.all(out[~scattered, 2:] == 1.0)
    # scatter distance from the entry: truncated exponential
    x = out[scattered, 1] - t0
    cdf = lambda v: (1 - np.exp(-mu_t * v)) / p  # noqa: E731
    assert stats.kstest(x, cdf).pvalue > 1e-3


def test_spectral_tracking_weights_are_unbiased():
    """chromatic homogeneous medium: per channel c the weighted histogram of
    scatter distances is mu_s,c exp(-mu_t,c t) and the weighted pass
    probability exp(-mu_t,c d) (Tracking::tracking's history-aware MIS)"""
    grid = np.full((8, 8, 8), 1.0, np.float32)
    cc_a = np.array([0.2, 1.0, 0.5])
    cc_s = np.array([1.5, 0.3, 0.9])
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
    n = 200000
    r = rays_through(n, (0.5, 0.5, -1.0), (0, 0, 1))
    out = m.batch(r, seed=11)
    t0 = float(np.float32(O.offset_f(1.0)))
    d = 2.0 - t0
    mu_t = cc_a + cc_s
    passed = out[:, 0] == 0
    for c in range(3):
        est = np.where(passed, out[:, 2 + c], 0.0)
        expect = np.exp(-mu_t[c] * d)
        assert abs(est.mean() - expect) < 5 * est.std() / np.sqrt(n) + 1e-4, (c, est.mean(), expect)
    edges = np.linspace(0, d, 6)
    sc = out[:, 0] == 1
    x = out[:, 1] - t0
    for c in range(3):
        for b in range(5):
            inb = sc & (x >= edges[b]) & (x < edges[b + 1])
            est = np.where(inb, out[:, 2 + c], 0.0)
            expect = cc_s[c] / mu_t[c] * (np.exp(-mu_t[c] * edges[b]) - np.exp(-mu_t[c] * edges[b + 1]))
            assert abs(est.mean() - expect) < 5 * est.std() / np.sqrt(n) + 1e-4, (c, b)


@pytest.mark.parametrize("g", [0.0, 0.6, -0.4, 0.9])
def test_henyey_greenstein(g):
    # normalisation over the sphere (midpoint rule in cos)
    cs = np.linspace(-1, 1, 20001)
    mid = 0.5 * (cs[1:] + cs[:-1])
    vals = np.array([O.phase_hg(float(c), g) for c in mid])
    integral = 2 * np.pi * np.sum(vals) * (cs[1] - cs[0])
    assert abs(integral - 1.0) < 1e-3
    # sampling: pdf is the phase of the sampled direction, mean cosine g
    rng = np.random.default_rng(5)
    wo = np.array([0.3, -0.5, 0.81], np.float32)
    wo /= np.linalg.norm(wo)
    cos_fwd = []
    for _ in range(20000):
        wi, pdf = O.sample_phase(wo, float(rng.random()), float(rng.random()), g)
        assert abs(np.linalg.norm(wi) - 1) < 1e-5
        assert abs(pdf - O.phase_hg(float(np.dot(wo, wi)), g)) <= 1e-3 * pdf + 1e-6
        cos_fwd.append(float(np.dot(-wo, wi)))
    assert abs(np.mean(cos_fwd) - g) < 5 * np.std(cos_fwd) / np.sqrt(len(cos_fwd)) + 1e-3


def test_similarity_scale():
    g = 0.6
    for depth in range(16):
        assert O.similarity_scale(g, depth) == 1.0
    for depth in (64, 100):
        assert abs(O.similarity_scale(g, depth) - (1 - g)) < 1e-7
    # halfway: gs = g / 2
    assert abs(O.similarity_scale(g, 40) - (1 - g) / (1 - g / 2)) < 1e-6


def cloud_box_scene(n=16):
    from sprout_amd import scenes

    pos, idx, tmat, mats = scenes.room()
    rng = np.random.default_rng(9)
    grid = (rng.random((n, n, n)) * 1.5).astype(np.float32)
    med = O.Medium(grid, (-0.6, 0.3, -0.6), (0.6, 1.5, 0.6), (0.3, 0.4, 0.5), (0.8, 0.7, 0.6), 0.5)
    return (pos, idx, tmat, mats), med


def test_medium_estimators_agree():
    """MIS, BSDF-only and NEE-only path tracing of the room with a medium
    converge to the same image"""
    from sprout_amd import scenes

    from scene_util import light_tree

    scene, med = cloud_box_scene()
    tree = light_tree(*scene)
    w, h, spp = 12, 12, 1024
    imgs = {}
    for mode in (O.PT_MIS, O.PT_BSDF, O.PT_NEE):
        imgs[mode] = O.render(*scene, w, h, spp, scenes.ROOM_CAMERA, tree, seed=5 + mode,
                              mode=mode, medium=med) / spp
    # compare means over 3 x 3 pixel blocks (less noise)
    def blocks(x):
        return x.reshape(4, 3, 4, 3, 3).mean(axis=(1, 3))
    ref = blocks(imgs[O.PT_MIS])
    for mode in (O.PT_BSDF, O.PT_NEE):
        rel = np.abs(blocks(imgs[mode]) - ref) / (ref + 0.02)
        assert np.median(rel) < 0.05, (mode, np.median(rel))
        assert abs(imgs[mode].mean() - imgs[O.PT_MIS].mean()) < 0.03 * imgs[O.PT_MIS].mean()
    # the medium attenuates: the image is darker than the empty room's
    empty = O.render(*scene, w, h, spp // 4, scenes.ROOM_CAMERA, tree, seed=5) / (spp // 4)
    assert imgs[O.PT_MIS].mean() < empty.mean()


def test_history_aware_scatter_probability():
    """In a homogeneous chromatic medium with no absorption of the majorant
    (ps + pn = 1) the k-th collision scatters with probability ps_k, where
    the channel weights follow lw_c <- lw_c mu_n,c / (mt pn) after every null
    collision and ps_k = avg(mu_s lw) / (avg(mu_s lw) + avg(mu_n lw)).  The
    number of collisions before d is Poisson(mt d), so P(scatter before d) =
    sum_k P(N >= k) ps_k prod_{j<k} (1 - ps_j).  Probabilities that ignore
    the weights' history (constant ps) give a different value."""
    grid = np.full((8, 8, 8), 1.0, np.float32)
    cc_a = np.array([0.0, 1.0, 1.0])
    cc_s = np.array([2.0, 0.1, 0.1])
    m = O.Medium(grid, (0, 0, 0), (1, 1, 1), cc_a, cc_s, 0.0)
    n = 200000
    out = m.batch(rays_through(n, (0.5, 0.5, -1.0), (0, 0, 1)), seed=13)
    t0 = O.offset_f(1.0)
    d = 2.0 - t0
    mt = cc_a.max() + cc_s.max()
    mu_n = mt - (cc_a + cc_s)
    lw = np.ones(3)
    p_none = 1.0  # probability that every collision so far was null
    total = 0.0
    const_total = 0.0
    ps0 = np.mean(cc_s) / (np.mean(cc_s) + np.mean(mu_n))
    for k in range(1, 200):
        ms, mn = np.mean(cc_s * lw), np.mean(mu_n * lw)
        ps = ms / (ms + mn)
        tail = stats.poisson.sf(k - 1, mt * d)  # P(N >= k)
        total += tail * p_none * ps
        const_total += tail * (1 - ps0) ** (k - 1) * ps0
        p_none *= 1 - ps
        lw = lw * mu_n / (mt * (1 - ps))
    got = np.mean(out[:, 0] == 1)
    sem = np.sqrt(total * (1 - total) / n)
    assert abs(got - total) < 5 * sem, (got, total)
    assert abs(const_total - total) > 20 * sem  # the check can tell the two apart


def test_medium_mis_weights_at_medium_vertices():
    """The cloud scene (ground, a large emissive quad above, a medium box in
    between): MIS and BSDF-only estimates agree.  The MIS weight of an
    emitter hit after a scattering vertex needs the light tree's pdf over
    the whole sphere; with the surface (normal-bounded) pdf the MIS image is
    brighter."""
    from sprout_amd import scenes

    from scene_util import light_tree

    scene = scenes.cloud_scene()
    tree = light_tree(*scene)
    rng = np.random.default_rng(21)
    grid = (rng.random((16, 16, 16)) * 2.0).astype(np.float32)
    med = O.Medium(grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S, scenes.CLOUD_G)
    w, h, spp = 8, 8, 512
    mis = O.render(*scene, w, h, spp, scenes.CLOUD_CAMERA, tree, seed=3, mode=O.PT_MIS,
                   medium=med, min_bounces=4, max_bounces=6) / spp
    bsdf = O.render(*scene, w, h, spp * 4, scenes.CLOUD_CAMERA, tree, seed=4, mode=O.PT_BSDF,
                    medium=med, min_bounces=4, max_bounces=6) / (4 * spp)
    assert abs(mis.mean() - bsdf.mean()) < 0.01 * bsdf.mean(), (mis.mean(), bsdf.mean())
