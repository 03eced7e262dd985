"""FP64 CPU oracle of the surrogate closure's multilayer perceptron.

TEST INFRASTRUCTURE ONLY (the checker of the BF16 MFMA kernels).  The network
has no counterpart in the reference; this restates it from its definition:
linear layers y = x W^T + b, GELU (tanh form) between them, a weighted
mean-squared-error loss sum(w (y - t)^2) / sum(w), and its analytic gradient
by reverse-mode differentiation.  ``bf16=True`` additionally rounds weights,
inputs, hidden activations and the back-propagated gradients to bfloat16
where the kernels store them, so that the GPU result can be checked tightly;
``bf16=False`` is the exact FP64 reference.
"""

from __future__ import annotations

import numpy as np

C_GELU = np.sqrt(2.0 / np.pi)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round to nearest-even bfloat16, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * x * x)


def unpack(widths, params):
    """Split the public parameter vector into [(W (out x in), b (out))]."""
    out, q = [], 0
    for i, o in zip(widths[:-1], widths[1:]):
        W = np.asarray(params[q:q + o * i], dtype=np.float64).reshape(o, i)
        q += o * i
        b = np.asarray(params[q:q + o], dtype=np.float64)
        q += o
        out.append((W, b))
    assert q == len(params)
    return out


def n_params(widths) -> int:
    return sum(o * i + o for i, o in zip(widths[:-1], widths[1:]))


def init_params(widths, seed=0) -> np.ndarray:
    """Scaled-uniform (Glorot) initialisation, deterministic in the seed."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in zip(widths[:-1], widths[1:]):
        lim = np.sqrt(6.0 / (i + o))
        parts.append(rng.uniform(-lim, lim, size=o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


def forward(widths, params, x, bf16=False):
    layers = unpack(widths, params)
    r = bf16_round if bf16 else (lambda a: a)
    h = r(np.asarray(x, dtype=np.float64))
    pres, acts = [], [h]
    for li, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        if li < len(layers) - 1:
            pres.append(z)
            h = r(gelu(z))
            acts.append(h)
        else:
            return z, pres, acts
    raise AssertionError


def loss_grad(widths, params, x, t, w, bf16=False):
    """Weighted MSE and its gradient in the public parameter layout."""
    layers = unpack(widths, params)
    r = bf16_round if bf16 else (lambda a: a)
    y, pres, acts = forward(widths, params, x, bf16)
    count = float(np.sum(w))
    e = y - t
    loss = float(np.sum(w * e * e) / count)
    # the output gradient 2 w e travels back unnormalised (as stored in
    # BF16); the 1 / count of the mean scales the summed parameter gradient
    d = r(2.0 * w * e)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        gW = (d.T @ acts[li]) / count
        gb = d.sum(axis=0) / count
        grads[li] = (gW, gb)
        if li > 0:
            dx = d @ r(W)
            d = r(dx * gelu_grad(pres[li - 1]))
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam_step(params, grad, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1) in FP64.

    ``m``, ``v`` are the first and second moment estimates before the step
    and ``t`` the 1-based step count; returns (params, m, v) after it:
    m <- beta1 m + (1 - beta1) g,  v <- beta2 v + (1 - beta2) g^2,
    p <- p - lr (m / (1 - beta1^t)) / (sqrt(v / (1 - beta2^t)) + eps)."""
    g = np.asarray(grad, dtype=np.float64)
    m = beta1 * np.asarray(m, dtype=np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1.0 - beta2) * g * g
    c1 = 1.0 - beta1 ** t
    c2 = 1.0 - beta2 ** t
    p = np.asarray(params, dtype=np.float64) - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v
