"""Pins the MLP oracle (oracle/mlp_oracle.py) the GPU training kernels are checked against.

  * exact float64 gradients equal central finite differences of the loss;
  * they equal torch autograd (CPU, float64) on the same network and loss;
  * the bfloat16-emulating gradients stay close to the exact ones, and the
    bf16 rounding itself is round-to-nearest-even (known bit patterns).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _batch(seed, B=16):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, 44)).astype(np.float32)
    Y = rng.standard_normal((B, 3))
    Y /= np.linalg.norm(Y, axis=1, keepdims=True)
    return X, Y


def test_bf16_rounding_known_bits():
    # bf16 keeps 7 mantissa bits: the spacing above 1 is 2^-7 = 0.0078125
    x = np.array([1.0, 1.00390625, 1.005859375, 1.001953125, 1.01171875, -2.5, 3.0e38], np.float32)
    r = mo.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 0.75*2^-7 rounds up; 1 + 2^-9 rounds down;
    # 1 + 1.5*2^-7 is a tie -> even (1 + 2^-6)
    assert list(r[:5]) == [1.0, 1.0, 1.0078125, 1.0, 1.015625]
    assert r[5] == -2.5 and np.isfinite(r[6])


def test_exact_grads_match_finite_differences():
    p = mo.he_init(1).astype(np.float64)
    X, Y = _batch(2)
    _, g = mo.loss_and_grads(p, X, Y, emulate=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(mo.NUM_PARAMS, 60, replace=False),
                          [0, 5631, 5632, 5759, mo.NUM_PARAMS - 4, mo.NUM_PARAMS - 1]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.loss_and_grads(pp, X, Y, emulate=False)[0] - mo.loss_and_grads(pm, X, Y, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_exact_grads_match_torch_autograd():
    torch = pytest.importorskip("torch")
    p = mo.he_init(4).astype(np.float64)
    X, Y = _batch(5, B=32)
    loss, g = mo.loss_and_grads(p, X, Y, emulate=False)
    tp = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(X.astype(np.float64))
    k = 0
    for li, (o, i) in enumerate(mo.LAYERS):
        W = tp[k:k + o * i].view(o, i)
        k += o * i
        b = tp[k:k + o]
        k += o
        h = h @ W.T + b
        if li < 3:
            h = torch.relu(h)
    y = torch.tensor(Y)
    tl = (1 - (h * y).sum(1) / h.norm(dim=1)).mean()
    tl.backward()
    assert abs(tl.item() - loss) <= 1e-12
    assert np.allclose(tp.grad.numpy(), g, rtol=1e-9, atol=1e-12)


def test_emulated_grads_close_to_exact():
    p = mo.he_init(6)
    X, Y = _batch(7, B=64)
    l0, g0 = mo.loss_and_grads(p, X, Y, emulate=False)
    l1, g1 = mo.loss_and_grads(p, X, Y, emulate=True)
    assert abs(l0 - l1) <= 2e-2 * max(1.0, abs(l0))
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    assert rel < 5e-2, rel
