"""The network oracle (oracle/mlp_oracle.py) pinned by finite differences and torch autograd."""

import numpy as np
import pytest

from oracle import mlp_oracle as M

DIMS = (7, 16, 3, 5)


def setup(seed=0, n=9):
    rng = np.random.default_rng(seed)
    theta = rng.normal(0, 0.4, M.param_count(*DIMS))
    x = rng.normal(0, 1, (n, DIMS[0]))
    y = rng.normal(0, 1, (n, DIMS[3]))
    w = rng.uniform(0.5, 2.0, DIMS[3])
    return theta, x, y, w


def test_backprop_matches_central_differences():
    theta, x, y, w = setup()
    _, g = M.loss_and_grad(theta, x, y, w, DIMS)
    h = 1e-6
    rng = np.random.default_rng(1)
    for k in rng.choice(len(theta), 60, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (M.loss_and_grad(tp, x, y, w, DIMS)[0] - M.loss_and_grad(tm, x, y, w, DIMS)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    theta, x, y, w = setup(2)
    loss, g = M.loss_and_grad(theta, x, y, w, DIMS)
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(x)
    k = 0
    for li, (o, i) in enumerate(M.shapes(*DIMS)):
        W = t[k:k + o * i].reshape(o, i)
        k += o * i
        b = t[k:k + o]
        k += o
        z = h @ W.T + b
        h = torch.nn.functional.gelu(z, approximate="tanh") if li < len(M.shapes(*DIMS)) - 1 else z
    tl = (torch.tensor(w) * (h - torch.tensor(y)) ** 2).sum() / (x.shape[0] * DIMS[3])
    tl.backward()
    assert float(tl) == pytest.approx(loss, rel=1e-12)
    assert np.allclose(t.grad.numpy(), g, rtol=1e-10, atol=1e-14)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 1e-30], np.float32)
    r = M.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                     # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even
    assert r[2] == 1.0                     # 1 + 2^-9 is below the midpoint
    assert r[3] == 1.0 + 2 ** -7           # 1 + 3 * 2^-9 is above the midpoint
    assert r[4] == -2.5
    assert np.all(np.abs(M.bf16_round(np.linspace(-3, 3, 1001)) - np.linspace(-3, 3, 1001)) <= 3 * 2 ** -8)


def test_adam_first_step_closed_form():
    p = np.array([0.5, -1.0, 2.0])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = M.adam_step(p, g, np.zeros(3), np.zeros(3), 1, 1e-3, 0.9, 0.999, 1e-8)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
