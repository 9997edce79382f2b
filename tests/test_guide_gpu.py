"""GPU guiding network (BF16 MFMA forward, gradients, Adam) against the FP64
oracle (oracle/guide.py).

The forward pass is compared twice: with the oracle emulating the kernel's
rounding (features, weights and hidden activations rounded to bf16, FP64
arithmetic otherwise), which pins the MFMA tiling, operand layouts and
layer order to FP32-accumulation noise; and with the plain FP64 oracle,
which bounds what BF16 costs.  Gradients are compared with the FP64
oracle per layer (direction and norm), Adam against the oracle's update of
the same gradient.
"""

import numpy as np
import pytest

from oracle import guide as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch

    from sprout_amd import bvh, guide

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    bvh.init(0)
    yield guide
    torch.cuda.synchronize()


def log(msg):
    print(msg, flush=True)


def bf16(a):
    """round to nearest even bfloat16, as float64"""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def emulated_forward(params, x):
    h = bf16(G.encode(x))
    for li, (w, b) in enumerate(params):
        z = h @ bf16(w) + np.asarray(b, np.float32).astype(np.float64)
        h = bf16(np.maximum(z, 0)) if li < 3 else z
    return h


def params_for_test(seed):
    p = G.init_params(seed)
    p[3][0] *= 3.0
    rng = np.random.default_rng(seed + 100)
    for w, b in p:
        b[:] = rng.normal(scale=0.1, size=b.shape)
    p[0][1][:] = np.abs(p[0][1])  # keep layer-1 units alive
    return p


def batch(n, seed):
    rng = np.random.default_rng(seed)
    # float32 inputs, as the device sees them: 2^11 pi x amplifies the
    # float64 -> float32 rounding of x to about 4e-4 in the top features
    x = rng.random((n, 9)).astype(np.float32).astype(np.float64)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = rng.exponential(size=n)
    return x, d, t


@pytest.fixture(scope="module")
def net(gpu):
    g = gpu.Guide(1)
    yield g
    g.release()


@pytest.mark.parametrize("n", [1, 15, 16, 1000, 40000])
def test_forward_matches_emulated_bf16(gpu, net, n):
    params = params_for_test(2)
    net.set_params(gpu.flatten(params))
    x, _, _ = batch(n, 3)
    log(f"forward n={n}")
    out = net.forward(x).cpu().numpy().astype(np.float64)
    ref = emulated_forward(params, x)
    # what remains is FP32 accumulation order: it moves a hidden activation
    # across a bf16 rounding boundary now and then (one bf16 ulp, 2^-8), and
    # the flips propagate; an operand-layout or tiling mistake gives errors of
    # the size of the outputs themselves
    diff = np.abs(out - ref)
    scale = np.abs(ref).max()
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1.5e-3
    assert np.percentile(diff, 99) / scale < 3e-3
    assert diff.max() / scale < 1e-2


def test_forward_close_to_fp64(gpu, net):
    params = params_for_test(4)
    net.set_params(gpu.flatten(params))
    x, _, _ = batch(4096, 5)
    out = net.forward(x).cpu().numpy().astype(np.float64)
    ref, _ = G.forward(params, G.encode(x))
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 3e-2, rel


def test_params_round_trip(gpu, net):
    p = gpu.flatten(params_for_test(6))
    net.set_params(p)
    assert np.array_equal(net.params(), p)


def emulated_backward(params, x, w, t):
    """G.backward with the kernel's roundings: features, weights, hidden
    activations and the back-propagated gradients (output gradient and
    each hidden layer's) rounded to bf16, FP64 otherwise."""
    feat = bf16(G.encode(x))
    acts = []
    h = feat
    for li, (wl, b) in enumerate(params):
        acts.append(h)
        z = h @ bf16(wl) + np.asarray(b, np.float32).astype(np.float64)
        h = bf16(np.maximum(z, 0)) if li < 3 else z
    dz = bf16(G.d_raw(h, w, t))
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        wl, _ = params[li]
        grads[li] = [acts[li].T @ dz, dz.sum(0)]
        if li > 0:
            dz = bf16((dz @ bf16(wl).T) * (acts[li] > 0))
    return grads


def compare(a, b):
    a = a.reshape(-1)
    b = b.reshape(-1)
    cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
    rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
    return cos, rel


@pytest.mark.parametrize("n", [64, 1000, 50000])
def test_gradients_match_emulated_bf16_and_fp64(gpu, net, n):
    """Against the oracle emulating the kernel's bf16 roundings the
    gradients agree to FP32-accumulation noise (a tiling, layout or
    reduction mistake moves them by their own size).  Against plain FP64
    they agree to what bf16 costs: a few samples with a small mixture
    density q carry weights t / q, and q depends exponentially on the
    rounded concentrations (at n = 1000 the emulated oracle itself is at
    cos 0.985, relative 0.17 from FP64 in the first layer)."""
    params = params_for_test(7)
    net.set_params(gpu.flatten(params))
    x, d, t = batch(n, 8)
    log(f"gradients n={n}")
    g = gpu.unflatten(net.gradients(x, d, t).cpu().numpy())
    assert np.all(g[0][0][G.NUM_FEAT:] == 0)
    emu = emulated_backward(params, x, d, t)
    ref = G.backward(params, G.encode(x), d, t)
    for li in range(4):
        for k in range(2):
            cos, rel = compare(g[li][k], emu[li][k])
            assert cos > 0.9995 and rel < 0.03, ("emulated", li, k, cos, rel)
            cos, rel = compare(g[li][k], ref[li][k])
            assert cos > 0.97 and rel < 0.25, ("fp64", li, k, cos, rel)


def test_gradients_are_deterministic(gpu, net):
    net.set_params(gpu.flatten(params_for_test(9)))
    x, d, t = batch(20000, 10)
    a = net.gradients(x, d, t).cpu().numpy()
    b = net.gradients(x, d, t).cpu().numpy()
    assert np.array_equal(a, b)


def test_adam_matches_oracle_update(gpu, net):
    params = params_for_test(11)
    flat = gpu.flatten(params)
    net.set_params(flat)
    x, d, t = batch(3000, 12)
    g = net.gradients(x, d, t).cpu().numpy().astype(np.float64)
    net.train_step(x, d, t, lr=2e-3)
    after = net.params().astype(np.float64)
    # the oracle's Adam on the same (GPU) gradient
    p = gpu.unflatten(flat.astype(np.float64))
    G.adam(p, gpu.unflatten(g), {}, lr=2e-3)
    expect = gpu.flatten(p).astype(np.float64)
    assert np.abs(after - expect).max() < 1e-6


def test_training_fits_a_lobe(gpu, net):
    """Online Adam on samples of one vMF lobe drives the loss toward it."""
    rng = np.random.default_rng(13)
    n = 8192
    mu = np.array([0.3, 0.4, np.sqrt(1 - 0.25)])
    d = np.array([G.sample_vmf(mu, 20.0, rng.random(), rng.random()) for _ in range(n)])
    x = np.tile(rng.random((1, 9)), (n, 1))
    t = np.ones(n)
    params = G.init_params(14)
    net.set_params(gpu.flatten(params))
    l0 = G.loss(params, G.encode(x[:256]), d[:256], t[:256])
    for _ in range(300):
        net.train_step(x, d, t, lr=3e-3)
    p = gpu.unflatten(net.params())
    l1 = G.loss(p, G.encode(x[:256]), d[:256], t[:256])
    # the loss cannot go below the entropy of the target lobe,
    # -(log C(k) + k A(k)) with C(k) = k / (4 pi sinh k), A(k) = coth k - 1 / k
    # (about -0.16 at k = 20): the fit must come close to it
    k = 20.0
    log_c = np.log(k) - np.log(4 * np.pi) - (k + np.log1p(-np.exp(-2 * k)) - np.log(2.0))
    entropy = -(log_c + k * (1 / np.tanh(k) - 1 / k))
    assert l1 < l0 - 2.0, (l0, l1)
    assert l1 < entropy + 0.15, (l1, entropy)
