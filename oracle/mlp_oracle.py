"""FP64 CPU restatement of the surrogate network, for checking the GPU kernels.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's checks import it; the product never does).
The surrogate has no counterpart in the reference; this module is pinned by central
finite differences (tests/test_mlp_oracle.py), not by reference vectors.

Networks: 5 -> 256 -> 256 -> 256 -> 256 -> 8 with SiLU after each hidden layer (the wide
network), and 5 -> 128 -> 128 -> 128 -> 128 -> 8 with hard SiLU, z clamp(z/6 + 1/2, 0, 1)
(the fast one).  Layer l maps x (B x n_in) to x @ W_l + b_l with W_l stored (n_in x n_out),
row-major.

Loss per ray (mean over the batch):
    cross-entropy of the 3 class logits (escaped, captured, disk)
    + sky weight * sum_k (y_{3+k} - t_k)^2,  k = 0..2                      escaped rays only
    + disk weight * [(y_6 - t_3)^2 + (y_7 - t_4)^2]                          disk rays only
Outputs 3..5 are the escape direction's components along the ray's straight-line frame,
6 and 7 are log r_hit and log g.
"""
from __future__ import annotations

import numpy as np

SIZES = (5, 256, 256, 256, 256, 8)
FAST_SIZES = (5, 128, 128, 128, 128, 8)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def hsilu(z):
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def dhsilu(z):
    return np.where(z <= -3.0, 0.0, np.where(z >= 3.0, 1.0, z / 3.0 + 0.5))


ACT = {"silu": (silu, dsilu), "hard": (hsilu, dhsilu)}


def init_params(seed: int = 0, sizes=SIZES):
    """He-uniform weights, zero biases (the GPU initialiser draws its own; tests copy)."""
    rng = np.random.default_rng(seed)
    params = []
    for n_in, n_out in zip(sizes[:-1], sizes[1:]):
        lim = np.sqrt(6.0 / n_in)
        params.append((rng.uniform(-lim, lim, (n_in, n_out)), np.zeros(n_out)))
    return params


def forward(params, x, act="silu"):
    """Returns the output (B x 8) and the per-layer (input, pre-activation) cache."""
    f = ACT[act][0]
    cache = []
    h = x
    for l, (W, b) in enumerate(params):
        z = h @ W + b
        cache.append((h, z))
        h = f(z) if l < len(params) - 1 else z
    return h, cache


def loss_and_grad_out(y, cls, tgt, w_sky=1.0, w_disk=1.0):
    """Loss (mean over the batch) and dL/dy for outputs y (B x 8), classes cls (B,),
    regression targets tgt (B x 5)."""
    B = y.shape[0]
    logits = y[:, :3]
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    p = e / e.sum(axis=1, keepdims=True)
    ce = -np.log(p[np.arange(B), cls])
    g = np.zeros_like(y)
    g[:, :3] = p
    g[np.arange(B), cls] -= 1.0
    sky = (cls == 0).astype(float)[:, None]
    disk = (cls == 2).astype(float)[:, None]
    d_sky = (y[:, 3:6] - tgt[:, 0:3]) * sky
    d_disk = (y[:, 6:8] - tgt[:, 3:5]) * disk
    loss = ce + w_sky * (d_sky ** 2).sum(axis=1) + w_disk * (d_disk ** 2).sum(axis=1)
    g[:, 3:6] = 2.0 * w_sky * d_sky
    g[:, 6:8] = 2.0 * w_disk * d_disk
    return loss.mean(), g / B


def backward(params, cache, gy, act="silu"):
    """Gradients [(dW, db)] for dL/dy = gy."""
    df = ACT[act][1]
    grads = [None] * len(params)
    g = gy
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        h, z = cache[l]
        if l < len(params) - 1:
            g = g * df(z)
        grads[l] = (h.T @ g, g.sum(axis=0))
        g = g @ W.T
    return grads


def loss_fn(params, x, cls, tgt, act="silu", **kw):
    y, _ = forward(params, x, act)
    return loss_and_grad_out(y, cls, tgt, **kw)[0]


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step on FP64 copies; state = {"t": int, "m": [...], "v": [...]}."""
    state["t"] += 1
    t = state["t"]
    out = []
    for i, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        new = []
        for j, (p, gp) in enumerate(((W, gW), (b, gb))):
            m = state["m"][2 * i + j] = beta1 * state["m"][2 * i + j] + (1 - beta1) * gp
            v = state["v"][2 * i + j] = beta2 * state["v"][2 * i + j] + (1 - beta2) * gp * gp
            mh = m / (1 - beta1 ** t)
            vh = v / (1 - beta2 ** t)
            new.append(p - lr * mh / (np.sqrt(vh) + eps))
        out.append(tuple(new))
    return out


def adam_state(params):
    zeros = []
    for W, b in params:
        zeros += [np.zeros_like(W), np.zeros_like(b)]
    return {"t": 0, "m": [z.copy() for z in zeros], "v": [z.copy() for z in zeros]}
