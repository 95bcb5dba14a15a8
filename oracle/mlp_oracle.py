"""ORACLE — test infrastructure, not product code.

Float64 forward and backward of the surrogate's network in numpy, the checker for the GPU's BF16
MFMA kernels (tissuemc_amd/csrc/mlp.hip). The network: inputs -> depth hidden layers with GELU in
its tanh form, 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) -> linear outputs; loss
sum_j w_j (y_pred - y)^2 / (batch * outputs) per head. Parameters per head: for each layer,
W [out x in] row-major, then b [out] (the C ABI's layout).

Pinned by central finite differences of its own loss (tests/test_mlp_oracle.py) and by torch's
float64 CPU autograd on the same network. With `bf16=True` the forward rounds weights, inputs and
hidden activations to bfloat16 (round to nearest even), as the GPU does, and accumulates in
float64.
"""

from __future__ import annotations

import numpy as np

C_GELU = np.sqrt(2.0 / np.pi)
A_GELU = 0.044715


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and back, as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + A_GELU * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + A_GELU * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * A_GELU * z * z)


def shapes(inputs: int, hidden: int, depth: int, outputs: int) -> list[tuple[int, int]]:
    """(out, in) of each layer."""
    return [(hidden, inputs)] + [(hidden, hidden)] * (depth - 1) + [(outputs, hidden)]


def param_count(inputs, hidden, depth, outputs) -> int:
    return sum(o * i + o for o, i in shapes(inputs, hidden, depth, outputs))


def unpack(theta: np.ndarray, inputs, hidden, depth, outputs):
    out, k = [], 0
    for o, i in shapes(inputs, hidden, depth, outputs):
        W = theta[k:k + o * i].reshape(o, i)
        k += o * i
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    return out


def forward(theta, x, dims, bf16=False):
    """Outputs and the cache for backward. dims = (inputs, hidden, depth, outputs)."""
    layers = unpack(np.asarray(theta, np.float64), *dims)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    h = r(x)
    cache = []
    for W, b in layers[:-1]:
        z = h @ r(W).T + b
        cache.append((h, z))
        h = r(gelu(z))
    W, b = layers[-1]
    y = h @ r(W).T + b
    cache.append((h, None))
    return y, cache


def loss_and_grad(theta, x, y, w, dims, bf16=False):
    """Loss and its gradient with respect to theta (float64)."""
    yp, cache = forward(theta, x, dims, bf16)
    n, nout = y.shape
    d = yp - y
    loss = float(np.sum(w[None, :] * d * d) / (n * nout))
    dy = 2.0 * w[None, :] * d / (n * nout)
    layers = unpack(np.asarray(theta, np.float64), *dims)
    grads = []
    g = dy
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        h_in, _ = cache[li]
        grads.append((g.T @ h_in, g.sum(axis=0)))
        if li > 0:
            dh = g @ W
            _, z_prev = cache[li - 1]
            g = dh * gelu_grad(z_prev)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def adam_step(p, g, m, v, t, lr, b1, b2, eps, wd=0.0):
    """One Adam update (Kingma & Ba 2015) with bias correction; returns (p, m, v)."""
    g = g + wd * p
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
