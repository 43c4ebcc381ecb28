"""The CPU FP64 oracle of the surrogate network (oracle/mlp_oracle.py) pinned
by mathematics and by independent implementations: its backward pass
against central finite differences of its loss, its forward pass against
torch's float64 Linear/ReLU stack, and its BF16 rounding against torch's
float32 -> bfloat16 conversion (round to nearest, ties to even)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def _net(rng, n_in=8, width=12, depth=3, n_out=10):
    layers = []
    for n, k in mo.shapes(n_in, width, depth, n_out):
        layers.append((rng.normal(0, 1.0 / np.sqrt(k), (n, k)), rng.normal(0, 0.1, n)))
    return layers


def test_backward_equals_finite_differences():
    rng = np.random.default_rng(3)
    layers = _net(rng)
    x = rng.normal(0, 1, (7, 8))
    t = rng.normal(0, 1, (7, 10))
    loss, grads = mo.loss_and_grad(layers, x, t)
    assert loss == pytest.approx(mo.loss_only(layers, x, t), rel=1e-14)
    flat_g = mo.pack(grads)
    picks = rng.choice(flat_g.size, 60, replace=False)
    # every layer's first weight and last bias too
    o = 0
    for w, b in layers:
        picks = np.append(picks, [o, o + w.size + b.size - 1])
        o += w.size + b.size
    fd = mo.finite_difference(layers, x, t, picks)
    err = np.abs(fd - flat_g[picks])
    assert err.max() <= 1e-7 * max(1.0, np.abs(flat_g).max()), err.max()
    # a wrong transposition or a dropped ReLU mask would fail this by orders of magnitude
    assert np.abs(flat_g[picks]).max() > 1e-3


def test_forward_equals_torch_float64():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    layers = _net(rng, n_in=5, width=16, depth=4, n_out=9)
    x = rng.normal(0, 1, (11, 5))
    mods = []
    for i, (w, b) in enumerate(layers):
        lin = torch.nn.Linear(w.shape[1], w.shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(w))
            lin.bias.copy_(torch.from_numpy(b))
        mods.append(lin)
        if i < len(layers) - 1:
            mods.append(torch.nn.ReLU())
    ref = torch.nn.Sequential(*mods)(torch.from_numpy(x)).detach().numpy()
    y, _ = mo.forward(layers, x)
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_bf16_rounding_equals_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(0, 1, 20000).astype(np.float32),
                        np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 3.0e-20, 65504.0])])
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16_round(v), ref)
    # ties go to the even mantissa
    assert mo.bf16_round(np.float32(1 + 2 ** -8)) == 1.0
    assert mo.bf16_round(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    np.testing.assert_array_equal(mo.from_bf16_bits(mo.bf16_bits(v)), ref)
