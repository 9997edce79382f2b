"""Pins the FP64 guiding-network oracle (oracle/guide.py).

No reference implementation exists, so the checks are the mathematics:
finite differences of the loss for every layer's weights and biases, the
normalization of the von Mises-Fisher lobes and their mixture, the vMF
sampler against its own density (mean cosine closed form), and Adam's first
step (a step of exactly lr against the gradient's sign).
"""

import numpy as np
import pytest

from oracle import guide as G


def batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, G.NUM_IN))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    t = rng.exponential(size=n)
    return G.encode(x), w, t


def test_parameter_count():
    assert G.num_params() == 224 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 32 + 32


def test_encoding():
    x = np.array([[0.25] * 9])
    f = G.encode(x)
    assert f.shape == (1, 224)
    assert np.all(f[:, 216:] == 0)
    # input 0 (position), multiplier 1: sin(pi/4), cos(pi/4); multiplier 2^(1/2)
    a = np.pi * 0.25 * np.float32(np.sqrt(2.0))
    assert np.allclose(f[0, 0:4], [np.sqrt(0.5), np.sqrt(0.5), np.sin(a), np.cos(a)])
    # input 0, multiplier 2^(2/2) = 2: sin(pi/2) = 1, cos = 0
    assert np.allclose(f[0, 4:6], [1.0, 0.0])
    # input 3 (normal), multiplier 2^(4/4) = 2 at j = 4
    assert np.allclose(f[0, 24 * 3 + 8:24 * 3 + 10], [1.0, 0.0])
    # an explicit table: the old octave ladder 2^j
    octaves = np.tile(2.0 ** np.arange(12), (9, 1))
    g = G.encode(x, octaves)
    assert np.allclose(g[0, 2:4], [1.0, 0.0])
    assert np.allclose(g[0, 24 * 8 + 2:24 * 8 + 4], [1.0, 0.0])


def relu_pattern(params, feat):
    """which hidden units are active, over the batch"""
    _, acts = G.forward(params, feat)
    return tuple(bytes(np.asarray(a > 0).reshape(-1)) for a in acts[1:-1])


@pytest.mark.parametrize("layer", [0, 1, 2, 3])
def test_gradients_match_finite_differences(layer):
    params = G.init_params(3)
    # larger output-layer weights so the lobes are not all alike
    params[3][0] *= 4
    feat, w, t = batch(16, 5)
    grads = G.backward(params, feat, w, t)
    rng = np.random.default_rng(layer)
    for k in range(2):
        p = params[layer][k]
        g = grads[layer][k]
        flat = p.reshape(-1)
        picks = rng.choice(flat.size, size=min(24, flat.size), replace=False)
        if layer == 0 and k == 0:
            # rows of the zero-padded features have no effect
            rows = np.arange(flat.size) // p.shape[1]
            picks = rng.choice(np.nonzero(rows < G.NUM_FEAT)[0], size=24, replace=False)
        checked = 0
        for i in picks:
            old = flat[i]
            h = 1e-6
            flat[i] = old + h
            lp = G.loss(params, feat, w, t)
            ap = relu_pattern(params, feat)
            flat[i] = old - h
            lm = G.loss(params, feat, w, t)
            am = relu_pattern(params, feat)
            flat[i] = old
            if ap != am:
                continue  # a ReLU switches inside [-h, h]: the difference quotient is not a derivative
            checked += 1
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g.reshape(-1)[i]) < 1e-6 + 1e-5 * abs(fd), (layer, k, i, fd, g.reshape(-1)[i])
        assert checked >= len(picks) - 2


def sphere_grid(n):
    """Midpoint rule on (cos theta, phi): equal-area cells."""
    z = (np.arange(n) + 0.5) / n * 2 - 1
    phi = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    zz, pp = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - zz * zz)
    w = np.stack([r * np.cos(pp), r * np.sin(pp), zz], -1).reshape(-1, 3)
    return w, 4 * np.pi / len(w)


def test_vmf_and_mixture_integrate_to_one():
    w, da = sphere_grid(800)
    for kappa in (0.05, 1.0, 20.0, 300.0):
        mu = np.array([[[0.0, 0.6, 0.8]]])
        v = G.vmf(np.repeat(mu, len(w), 0), np.full((len(w), 1), kappa), w)
        assert abs(v.sum() * da - 1) < 2e-3, kappa
    rng = np.random.default_rng(1)
    raw = rng.normal(size=(1, 32)) * 2
    q = G.mixture_pdf(np.repeat(raw, len(w), 0), w)
    assert abs(q.sum() * da - 1) < 2e-3


@pytest.mark.parametrize("g", [-0.3, 0.0, 0.6, 0.9])
def test_medium_vertex_mixture_integrates_to_one(g):
    """At a scattering vertex the guided pdf is fraction * q(w) + (1 -
    fraction) * HG(wo.w, g) over the whole sphere; it integrates to one, and
    the one-sample mixture sampler (guide with probability fraction, else the
    phase function) draws directions with that density (binned in the
    cosine to wo)."""
    from oracle import oracle as O

    w, da = sphere_grid(800)
    rng = np.random.default_rng(2)
    raw = rng.normal(size=(1, 32)) * 2
    wo = np.array([0.48, -0.6, 0.64])
    fraction = 0.7
    q = G.mixture_pdf(np.repeat(raw, len(w), 0), w)
    ph = np.array([O.phase_hg(c, g) for c in w @ wo])
    assert abs(ph.sum() * da - 1) < 2e-3
    pdf = fraction * q + (1 - fraction) * ph
    assert abs(pdf.sum() * da - 1) < 2e-3

    mu, kappa, wts = (a[0] for a in G.lobes(raw))
    n = 20000
    samples = np.zeros((n, 3))
    for i in range(n):
        if rng.random() < fraction:
            lobe = min(np.searchsorted(np.cumsum(wts), rng.random()), len(wts) - 1)
            samples[i] = G.sample_vmf(mu[lobe], kappa[lobe], rng.random(), rng.random())
        else:
            samples[i], _ = O.sample_phase(wo, rng.random(), rng.random(), g)
    edges = np.linspace(-1, 1, 11)
    got = np.histogram(samples @ wo, edges)[0] / n
    want = np.array([pdf[(w @ wo >= a) & (w @ wo < b)].sum() * da for a, b in zip(edges, edges[1:])])
    assert np.all(np.abs(got - want) < 5 * np.sqrt(want * (1 - want) / n) + 2e-3), (got, want)


def test_vmf_sampler_mean_cosine():
    """E[mu.w] = coth(kappa) - 1/kappa for a vMF on the sphere."""
    rng = np.random.default_rng(4)
    mu = np.array([0.3, -0.5, 0.812403840463596])
    mu /= np.linalg.norm(mu)
    for kappa in (0.5, 5.0, 50.0):
        s = np.array([G.sample_vmf(mu, kappa, rng.random(), rng.random()) for _ in range(20000)])
        assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-9)
        expect = 1 / np.tanh(kappa) - 1 / kappa
        got = (s @ mu).mean()
        assert abs(got - expect) < 4 * (s @ mu).std() / np.sqrt(len(s)) + 1e-4, (kappa, got, expect)


def test_adam_first_step():
    params = G.init_params(0)
    feat, w, t = batch(8, 1)
    grads = G.backward(params, feat, w, t)
    before = [[a.copy() for a in p] for p in params]
    G.adam(params, grads, {}, lr=1e-3)
    for li in range(4):
        for k in range(2):
            step = before[li][k] - params[li][k]
            g = grads[li][k]
            big = np.abs(g) > 1e-3  # eps / |g| below 1e-5
            assert np.allclose(step[big], 1e-3 * np.sign(g[big]), rtol=1e-4)


def test_training_reduces_loss_toward_a_fixed_lobe():
    """A few hundred Adam steps on samples from one vMF lobe lower the loss."""
    rng = np.random.default_rng(7)
    mu = np.array([0.0, 0.0, 1.0])
    n = 256
    w = np.array([G.sample_vmf(mu, 10.0, rng.random(), rng.random()) for _ in range(n)])
    x = np.tile(rng.random((1, 9)), (n, 1))
    feat = G.encode(x)
    t = np.ones(n)
    params = G.init_params(2)
    state = {}
    l0 = G.loss(params, feat, w, t)
    for _ in range(200):
        G.adam(params, G.backward(params, feat, w, t), state, lr=3e-3)
    l1 = G.loss(params, feat, w, t)
    # the entropy of vMF(kappa = 10) is about -0.5: the fit gets close
    assert l1 < l0 - 1.0, (l0, l1)
