"""TEST INFRASTRUCTURE ONLY: CPU restatement of the learned GJK warm-start MLP.

The network has no counterpart in parry; it fills the ``init_dir`` slot of
contact_support_map_support_map_with_params
(src/query/contact/contact_support_map_support_map.rs:52-60).  This file is the
oracle the GPU forward/backward kernels are checked against; only tests/,
smoke() and bench.py load it.

Network: 44 -> 128 -> 128 -> 128 -> 3, ReLU on the hidden layers, linear
output.  Parameters are one flat float32 vector in the order
W1 (128x44, row-major), b1, W2 (128x128), b2, W3, b3, W4 (3x128), b4.

Loss (per batch of B pairs): mean over pairs of 1 - cos(out, label), label a
unit vector.  dL/dout = -(y/|o| - (o.y) o/|o|^3) / B.

``emulate=True`` reproduces where the GPU kernels round to bfloat16 (round to
nearest even): the weights, the input features, every hidden activation after
ReLU and every back-propagated delta (dL/dz) before it is used; products are
accumulated in float64 here (float32 on the GPU).  ``emulate=False`` is the
exact float64 network, pinned by finite differences and torch autograd.
"""
from __future__ import annotations

import numpy as np

LAYERS = ((128, 44), (128, 128), (128, 128), (3, 128))
NUM_PARAMS = sum(o * i + o for o, i in LAYERS)
assert NUM_PARAMS == 39171


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(a)
    r = r.astype(np.uint32)
    r[nan] = 0x7FC00000
    return r.view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    assert p.shape == (NUM_PARAMS,)
    out, k = [], 0
    for o, i in LAYERS:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def _round(x, emulate):
    return bf16(x.astype(np.float32)) if emulate else x


def forward(params, X, emulate: bool = True):
    """Returns (out [B,3], list of layer inputs [X, h1, h2, h3] as used)."""
    layers = unpack(params)
    if emulate:
        layers = [(bf16(W), b.astype(np.float32).astype(np.float64)) for W, b in layers]
    h = _round(np.asarray(X, np.float64), emulate)
    acts = [h]
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        if li < 3:
            h = _round(np.maximum(z, 0.0), emulate)
            acts.append(h)
        else:
            h = z.astype(np.float32).astype(np.float64) if emulate else z
    return h, acts


def loss_grad_out(out, Y):
    """Mean 1 - cos loss and dL/dout (float64)."""
    B = out.shape[0]
    n = np.sqrt((out * out).sum(1, keepdims=True))
    dot = (out * Y).sum(1, keepdims=True)
    ok = n > 1e-12
    safe = np.where(ok, n, 1.0)
    cos = np.where(ok, dot / safe, 0.0)
    loss = float((1.0 - cos).sum() / B)
    g = -(Y / safe - dot * out / safe ** 3) / B
    g = np.where(ok, g, 0.0)
    return loss, g


def loss_and_grads(params, X, Y, emulate: bool = True):
    """(loss, flat gradient in the parameter order) for one batch."""
    layers = unpack(params)
    if emulate:
        layers = [(bf16(W), b) for W, b in layers]
    out, acts = forward(params, X, emulate)
    loss, d = loss_grad_out(out, np.asarray(Y, np.float64))
    if emulate:
        d = bf16(d.astype(np.float32))
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        h_in = acts[li]
        grads[li] = (d.T @ h_in, d.sum(0))
        if li == 0:
            break
        dh = d @ W
        d = dh * (acts[li] > 0)
        if emulate:
            d = bf16(d.astype(np.float32))
    return loss, pack(grads)


def he_init(seed: int) -> np.ndarray:
    """A float32 parameter vector with W ~ N(0, 2/fan_in), b = 0 (for tests)."""
    rng = np.random.default_rng(seed)
    return pack([(rng.standard_normal((o, i)) * np.sqrt(2.0 / i), np.zeros(o)) for o, i in LAYERS]).astype(np.float32)
