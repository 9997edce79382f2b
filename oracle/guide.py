"""ORACLE (test infrastructure only; never imported by the product).

FP64 restatement of the path-guiding network the GPU trains and queries:
the reference has no learned guiding (its nearest relatives are the
particle-importance maps of core/rendering/integrator/particle/
particle_importance.hpp), so this file is the definition the HIP kernels are
checked against, itself checked by finite differences and by the closed
forms of the von Mises-Fisher distribution (tests/test_oracle_guide.py).

Model
  inputs   x (9): position normalized to the scene box [0,1]^3, normal and
           outgoing direction mapped from [-1,1] to [0,1]
  features 12 multipliers f per input: sin(pi f x), cos(pi f x), f x in
           float32; feature 24 i + 2 j (+1 for cos); 216 features,
           zero-padded to 224; by default f = 2^(j/2) for the position and
           2^(j/4) for the normal and direction
  network  224 -> 64 -> 64 -> 64 (ReLU) -> 32, biases everywhere
  outputs  8 lobes x (a, b, c, d):
             theta = pi sigmoid(a), phi = 2 pi sigmoid(b)
             mu    = (sin theta cos phi, sin theta sin phi, cos theta)
             kappa = exp(clamp(c, -4, 8))
             pi_k  = softmax(d)_k
  density  q(w) = sum_k pi_k vMF(w; mu_k, kappa_k),
           vMF(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
  loss     per sample with weight t >= 0 and direction w:  -t log q(w),
           averaged over the batch (weights normalized by the batch mean)
"""

from __future__ import annotations

import numpy as np

NUM_FREQ = 12
NUM_IN = 9
NUM_FEAT = NUM_IN * NUM_FREQ * 2  # 216
FEAT_PAD = 224
HIDDEN = 64
NUM_LOBES = 8
NUM_OUT = 4 * NUM_LOBES
KAPPA_MIN, KAPPA_MAX = -4.0, 8.0

SHAPES = [(FEAT_PAD, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, NUM_OUT)]


def num_params() -> int:
    return sum(i * o + o for i, o in SHAPES)


def init_params(seed: int = 0) -> list:
    """He-uniform weights (rows of padded features zero), zero biases."""
    rng = np.random.default_rng(seed)
    params = []
    for li, (i, o) in enumerate(SHAPES):
        fan_in = NUM_FEAT if li == 0 else i
        bound = np.sqrt(6.0 / fan_in)
        w = rng.uniform(-bound, bound, (i, o))
        if li == 0:
            w[NUM_FEAT:] = 0.0
        params.append([w, np.zeros(o)])
    return params


def inputs(p, n, wo, box_lo, box_hi) -> np.ndarray:
    p = (np.asarray(p, np.float64) - box_lo) / (np.asarray(box_hi) - box_lo)
    return np.concatenate([p, (np.asarray(n) + 1) * 0.5, (np.asarray(wo) + 1) * 0.5], axis=-1)


def default_frequencies() -> np.ndarray:
    """(9, 12) float32 multipliers: 2^(j/2) for the position inputs, 2^(j/4)
    for the normal and direction inputs"""
    j = np.arange(NUM_FREQ, dtype=np.float64)
    step = np.array([0.5] * 3 + [0.25] * 6)[:, None]
    return np.exp2(step * j[None, :]).astype(np.float32)


def encode(x: np.ndarray, freq: np.ndarray | None = None) -> np.ndarray:
    """(B, 9) in [0,1] -> (B, 224): sin / cos of pi f x per input and
    multiplier, the product f x rounded to float32 as the kernels compute it"""
    b = x.shape[0]
    f = np.zeros((b, FEAT_PAD))
    freq = default_frequencies() if freq is None else np.asarray(freq, np.float32).reshape(9, NUM_FREQ)
    prod = (np.asarray(x, np.float32)[:, :, None] * freq[None, :, :]).astype(np.float64)
    arg = prod * np.pi  # (B, 9, 12)
    enc = np.stack([np.sin(arg), np.cos(arg)], axis=-1)  # (B, 9, 12, 2)
    f[:, :NUM_FEAT] = enc.reshape(b, NUM_FEAT)
    return f


def forward(params, feat):
    """Returns raw outputs (B, 32) and the activations for backward."""
    acts = [feat]
    h = feat
    for li, (w, bias) in enumerate(params):
        z = h @ w + bias
        if li < len(params) - 1:
            h = np.maximum(z, 0.0)
        else:
            h = z
        acts.append(h)
    return h, acts


def lobes(raw):
    """raw (B, 32) -> mu (B, 8, 3), kappa (B, 8), weights (B, 8)"""
    r = raw.reshape(-1, NUM_LOBES, 4)
    sa = 1 / (1 + np.exp(-r[..., 0]))
    sb = 1 / (1 + np.exp(-r[..., 1]))
    theta = np.pi * sa
    phi = 2 * np.pi * sb
    mu = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    kappa = np.exp(np.clip(r[..., 2], KAPPA_MIN, KAPPA_MAX))
    d = r[..., 3] - r[..., 3].max(axis=1, keepdims=True)
    e = np.exp(d)
    wts = e / e.sum(axis=1, keepdims=True)
    return mu, kappa, wts


def vmf(mu, kappa, w):
    """vMF density of directions w (B, 3) under lobes mu (B, L, 3), kappa (B, L)."""
    c = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa)))
    return c * np.exp(kappa * ((mu * w[:, None, :]).sum(-1) - 1))


def mixture_pdf(raw, w):
    mu, kappa, wts = lobes(raw)
    return (wts * vmf(mu, kappa, w)).sum(1)


def loss(params, feat, w, t):
    raw, _ = forward(params, feat)
    q = mixture_pdf(raw, w)
    tn = t / t.mean()
    return float(np.mean(-tn * np.log(q)))


def d_raw(raw, w, t):
    """dLoss/draw (B, 32) for loss = mean(-tn log q)."""
    b = raw.shape[0]
    r = raw.reshape(b, NUM_LOBES, 4)
    mu, kappa, wts = lobes(raw)
    v = vmf(mu, kappa, w)  # (B, L)
    q = (wts * v).sum(1)  # (B,)
    tn = t / t.mean()
    g = -(tn / b) / q  # dL/dq (B,)
    resp = wts * v  # pi_k vMF_k
    out = np.zeros_like(r)
    # softmax weights: dq/dd_j = pi_j (v_j - q)
    out[..., 3] = g[:, None] * wts * (v - q[:, None])
    # log vMF_k = log kappa - log(2pi) - log(1 - e^{-2k}) + kappa (mu.w - 1)
    cosw = (mu * w[:, None, :]).sum(-1)
    e2 = np.exp(-2 * kappa)
    dlogv_dkappa = 1 / kappa - 2 * e2 / (1 - e2) + (cosw - 1)
    clip_ok = (r[..., 2] > KAPPA_MIN) & (r[..., 2] < KAPPA_MAX)
    out[..., 2] = g[:, None] * resp * dlogv_dkappa * kappa * clip_ok
    # mu(theta, phi): d log v / d mu = kappa w
    sa = 1 / (1 + np.exp(-r[..., 0]))
    sb = 1 / (1 + np.exp(-r[..., 1]))
    theta = np.pi * sa
    phi = 2 * np.pi * sb
    dmu_dtheta = np.stack([np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi),
                           -np.sin(theta)], -1)
    dmu_dphi = np.stack([-np.sin(theta) * np.sin(phi), np.sin(theta) * np.cos(phi),
                         np.zeros_like(theta)], -1)
    kw = kappa[..., None] * w[:, None, :]
    dl_dtheta = (kw * dmu_dtheta).sum(-1)
    dl_dphi = (kw * dmu_dphi).sum(-1)
    out[..., 0] = g[:, None] * resp * dl_dtheta * np.pi * sa * (1 - sa)
    out[..., 1] = g[:, None] * resp * dl_dphi * 2 * np.pi * sb * (1 - sb)
    return out.reshape(b, NUM_OUT)


def backward(params, feat, w, t):
    """Gradients of loss() with respect to every parameter, same structure as params."""
    raw, acts = forward(params, feat)
    dz = d_raw(raw, w, t)
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        wl, _ = params[li]
        grads[li] = [acts[li].T @ dz, dz.sum(0)]
        if li > 0:
            dh = dz @ wl.T
            dz = dh * (acts[li] > 0)
    return grads


def sample_vmf(mu, kappa, u1, u2):
    """Direction from one vMF lobe (mu (3,), kappa) with uniforms u1, u2."""
    wz = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    r = np.sqrt(max(0.0, 1 - wz * wz))
    phi = 2 * np.pi * u2
    local = np.array([r * np.cos(phi), r * np.sin(phi), wz])
    # frame around mu (orthonormal_basis, base/math/vector3.inl:523-534)
    sign = 1.0 if mu[2] >= 0 else -1.0
    c = -1 / (sign + mu[2])
    d = mu[0] * mu[1] * c
    t = np.array([1 + sign * mu[0] * mu[0] * c, sign * d, -sign * mu[0]])
    b = np.array([d, sign + mu[1] * mu[1] * c, -mu[1]])
    return local[0] * t + local[1] * b + local[2] * mu


def adam(params, grads, state, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step in place; state = {'t': int, 'm': [...], 'v': [...]}."""
    if not state:
        state["t"] = 0
        state["m"] = [[np.zeros_like(a) for a in p] for p in params]
        state["v"] = [[np.zeros_like(a) for a in p] for p in params]
    state["t"] += 1
    t = state["t"]
    for li, (p, g) in enumerate(zip(params, grads)):
        for k in range(2):
            m = state["m"][li][k] = b1 * state["m"][li][k] + (1 - b1) * g[k]
            v = state["v"][li][k] = b2 * state["v"][li][k] + (1 - b2) * g[k] * g[k]
            mh = m / (1 - b1 ** t)
            vh = v / (1 - b2 ** t)
            p[k] -= lr * mh / (np.sqrt(vh) + eps)
