"""Float64 NumPy reference of the surrogate MLP.  TEST INFRASTRUCTURE ONLY:
the checker for the HIP forward / backward / Adam kernels.

Network 5 -> 64 -> 64 -> 64 -> 6, SiLU on the hidden layers, linear output,
loss = mean over the batch and the 6 outputs of (y - target)^2.  Parameters
are one flat vector: W1 (64x5, row-major out x in), b1, W2 (64x64), b2,
W3 (64x64), b3, W4 (6x64), b4 -- 9094 values.  The backward pass is the
analytic chain rule; tests pin it by central finite differences.
"""

from __future__ import annotations

import numpy as np

LAYERS = ((5, 64), (64, 64), (64, 64), (64, 6))
PARAM_COUNT = sum(i * o + o for i, o in LAYERS)


def unpack(flat: np.ndarray):
    flat = np.asarray(flat, dtype=np.float64)
    assert flat.shape == (PARAM_COUNT,)
    out, off = [], 0
    for fan_in, fan_out in LAYERS:
        w = flat[off:off + fan_in * fan_out].reshape(fan_out, fan_in)
        off += fan_in * fan_out
        b = flat[off:off + fan_out]
        off += fan_out
        out.append((w, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(flat, x):
    """Returns (output, cache of pre-activations and layer inputs)."""
    layers = unpack(flat)
    a = np.asarray(x, dtype=np.float64)
    inputs, pre = [], []
    for i, (w, b) in enumerate(layers):
        inputs.append(a)
        z = a @ w.T + b
        pre.append(z)
        a = silu(z) if i < len(layers) - 1 else z
    return a, (inputs, pre)


def loss(y, target) -> float:
    d = np.asarray(y, dtype=np.float64) - np.asarray(target, dtype=np.float64)
    return float(np.mean(d * d))


def gradients(flat, x, target):
    """d loss / d params as a flat vector, plus the loss."""
    layers = unpack(flat)
    y, (inputs, pre) = forward(flat, x)
    n = y.shape[0] * y.shape[1]
    dz = 2.0 * (y - np.asarray(target, dtype=np.float64)) / n
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        w, _ = layers[i]
        grads[i] = (dz.T @ inputs[i], dz.sum(axis=0))
        if i > 0:
            dz = (dz @ w) * dsilu(pre[i - 1])
    flat_grad = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return flat_grad, loss(y, target)


def adam_step(flat, grad, m, v, step, lr=1.0e-3, beta1=0.9, beta2=0.999, eps=1.0e-8):
    """One Adam update (Kingma & Ba, 2015, Algorithm 1); `step` counts from 1."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    m_hat = m / (1.0 - beta1 ** step)
    v_hat = v / (1.0 - beta2 ** step)
    return flat - lr * m_hat / (np.sqrt(v_hat) + eps), m, v


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases."""
    g = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in LAYERS:
        bound = np.sqrt(6.0 / fan_in)
        parts.append(g.uniform(-bound, bound, size=fan_in * fan_out))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts)


__all__ = ["LAYERS", "PARAM_COUNT", "adam_step", "dsilu", "forward", "gradients", "init_params",
           "loss", "silu", "unpack"]
