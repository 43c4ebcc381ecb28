"""CPU FP64 oracle of the flux surrogate's network (TEST INFRASTRUCTURE ONLY:
imported by tests/ and nothing else; the library never calls it).

The network is the one soltrace_amd/csrc/st_mlp.hip trains on the GPU:

    h_0 = x,  h_l = relu(W_l h_{l-1} + b_l)  (l = 1..depth),  y = W_out h_depth + b_out
    loss = mean over the batch and the outputs of (y - t)^2

Parameters travel in the library's exchange order: per layer W [n_out][n_in]
row-major, then b [n_out].  The reference has no such network (its flux
maps come from tracing alone, app/src/project.cpp:1139-1436), so this
oracle is pinned by the mathematics instead: central finite differences of
the FP64 loss against the closed-form backward pass (tests/test_mlp_oracle.py).

Two modes:
  exact    FP64 throughout (given the BF16 weights the GPU multiplies with);
  emulate  the same, with the values the GPU stores in BF16 rounded to BF16
           at the points it stores them: inputs, hidden activations and the
           backward pass's pre-activation gradients.  Only the order and
           width of accumulation then differ from the GPU.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a) -> np.ndarray:
    """Round to the nearest BF16 value, ties to even (finite inputs)."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(a) -> np.ndarray:
    """BF16 bit patterns (uint16) of the values, rounded as bf16_round."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def shapes(n_in: int, width: int, depth: int, n_out: int):
    """(n_out, n_in) of every linear layer."""
    dims = [n_in] + [width] * depth + [n_out]
    return [(dims[i + 1], dims[i]) for i in range(depth + 1)]


def unpack(flat, n_in, width, depth, n_out):
    """Exchange-order vector -> [(W, b), ...] in FP64."""
    flat = np.asarray(flat, dtype=np.float64)
    out, o = [], 0
    for n, k in shapes(n_in, width, depth, n_out):
        w = flat[o:o + n * k].reshape(n, k)
        o += n * k
        b = flat[o:o + n]
        o += n
        out.append((w, b))
    assert o == flat.size, (o, flat.size)
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers])


def forward(layers, x, emulate: bool = False):
    """Outputs y (B, n_out) and the activations [h_0 .. h_depth]."""
    r = bf16_round if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    h = [r(x)]
    for w, b in layers[:-1]:
        h.append(r(np.maximum(h[-1] @ w.T + b, 0.0)))
    w, b = layers[-1]
    return h[-1] @ w.T + b, h


def loss_and_grad(layers, x, t, emulate: bool = False):
    """Mean squared error and its gradient [(dW, db), ...] by backpropagation."""
    y, h = forward(layers, x, emulate)
    t = np.asarray(t, dtype=np.float64)
    bsz, d = y.shape
    diff = y - t
    loss = float(np.mean(diff * diff))
    dz = 2.0 * diff / (bsz * d)
    if emulate:
        dz = bf16_round(dz)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        w, _ = layers[li]
        grads[li] = (dz.T @ h[li], dz.sum(axis=0))
        if li > 0:
            dh = (dz @ w) * (h[li] > 0.0)
            dz = bf16_round(dh) if emulate else dh
    return loss, grads


def loss_only(layers, x, t) -> float:
    y, _ = forward(layers, x, False)
    d = y - np.asarray(t, dtype=np.float64)
    return float(np.mean(d * d))


def finite_difference(layers, x, t, picks, eps: float = 1e-6):
    """Central differences of the FP64 loss at parameter positions `picks`
    (indices into the exchange-order vector)."""
    flat = pack(layers)
    dims = [(w.shape[1], w.shape[0]) for w, _ in layers]
    n_in, width, depth, n_out = dims[0][0], dims[0][1], len(layers) - 1, dims[-1][1]
    out = []
    for i in picks:
        p = flat.copy()
        p[i] += eps
        lp = loss_only(unpack(p, n_in, width, depth, n_out), x, t)
        p[i] -= 2 * eps
        lm = loss_only(unpack(p, n_in, width, depth, n_out), x, t)
        out.append((lp - lm) / (2 * eps))
    return np.array(out)
