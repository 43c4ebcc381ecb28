"""The surrogate network's MFMA kernels (soltrace_amd/csrc/st_mlp.hip) against
the CPU FP64 oracle (oracle/mlp_oracle.py), which finite differences pin.

* forward: outputs equal the oracle's with the BF16 roundings emulated,
  to FP32 accumulation error, at the flux surrogate's shape (8 -> 6 x 256
  -> 4096) and a batch that is not a multiple of the 64-row tile;
* gradients of every tensor: against the emulating oracle to 1%, against
  the plain FP64 oracle (no BF16 activations) by cosine and norm;
* training: Adam steps drive the loss of a smooth regression down, and
  the exchange of parameters round-trips."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPE = (8, 256, 6, 4096)


def _mlp(hip_lib, shape=SHAPE, max_batch=512, seed=3):
    from soltrace_amd.surrogate import Mlp
    return Mlp(hip_lib, *shape, max_batch=max_batch, seed=seed)


def _data(rng, b, shape=SHAPE):
    import torch
    n_in, n_out = shape[0], shape[3]
    x = rng.uniform(-1, 1, (b, n_in)).astype(np.float32)
    t = rng.normal(0, 0.5, (b, n_out)).astype(np.float32)
    tb = mo.bf16_bits(t)
    xd = torch.from_numpy(x).cuda()
    td = torch.from_numpy(tb.view(np.int16)).cuda().view(torch.bfloat16)
    return x, mo.from_bf16_bits(tb), xd, td


def test_forward_equals_oracle(hip_lib):
    import torch
    net = _mlp(hip_lib)
    rng = np.random.default_rng(1)
    x, _, xd, _ = _data(rng, 100)
    layers = mo.unpack(net.get_params(1), *SHAPE)
    y = net.forward(xd)
    torch.cuda.synchronize()
    y = y.cpu().numpy().astype(np.float64)
    ye, _ = mo.forward(layers, x, emulate=True)
    scale = np.abs(ye).max()
    assert scale > 0.1
    err = np.abs(y - ye)
    print(f"forward: max |dy| / max |y| = {err.max() / scale:.2e}, median {np.median(err) / scale:.2e}")
    assert err.max() <= 2e-3 * scale and np.median(err) <= 2e-5 * scale
    # the plain FP64 network differs only by the BF16 activations
    yx, _ = mo.forward(layers, x, emulate=False)
    assert np.abs(y - yx).max() <= 3e-2 * scale
    net.close()


@pytest.mark.parametrize("batch", [64, 300])
def test_gradients_equal_oracle(hip_lib, batch):
    net = _mlp(hip_lib)
    rng = np.random.default_rng(batch)
    x, t, xd, td = _data(rng, batch)
    loss, g = net.gradients(xd, td)
    layers = mo.unpack(net.get_params(1), *SHAPE)
    le, ge = mo.loss_and_grad(layers, x, t, emulate=True)
    lx, gx = mo.loss_and_grad(layers, x, t, emulate=False)
    assert loss == pytest.approx(le, rel=2e-3)
    gg = mo.unpack(g.astype(np.float64), *SHAPE)
    for li, ((dw, db), (ew, eb), (xw, xb)) in enumerate(zip(gg, ge, gx)):
        for name, a, e, xx in (("W", dw, ew, xw), ("b", db, eb, xb)):
            rel = np.linalg.norm(a - e) / np.linalg.norm(e)
            cos = float((a * xx).sum() / (np.linalg.norm(a) * np.linalg.norm(xx)))
            relx = np.linalg.norm(a - xx) / np.linalg.norm(xx)
            print(f"layer {li} {name}: |g-g_emul|/|g_emul| {rel:.2e}, |g-g64|/|g64| {relx:.2e}, cos {cos:.6f}")
            assert rel <= 1e-2, (li, name, rel)
            assert cos >= 0.995 and relx <= 0.1, (li, name, cos, relx)
    net.close()


def test_training_reduces_loss_and_params_round_trip(hip_lib):
    import torch
    shape = (4, 128, 3, 256)
    net = _mlp(hip_lib, shape=shape, max_batch=256, seed=9)
    rng = np.random.default_rng(2)
    # a smooth map: a Gaussian spot on a 16 x 16 grid whose centre and width follow the inputs
    def batch(b):
        x = rng.uniform(-1, 1, (b, 4)).astype(np.float32)
        gx, gy = np.meshgrid(np.linspace(-1, 1, 16), np.linspace(-1, 1, 16), indexing="ij")
        s = 0.3 + 0.1 * x[:, 2:3, None]
        img = np.exp(-((gx[None] - 0.6 * x[:, 0:1, None]) ** 2 + (gy[None] - 0.6 * x[:, 1:2, None]) ** 2)
                     / (2 * s * s)) * (1 + 0.3 * x[:, 3:4, None])
        t = img.reshape(b, 256).astype(np.float32)
        return (torch.from_numpy(x).cuda(),
                torch.from_numpy(mo.bf16_bits(t).view(np.int16)).cuda().view(torch.bfloat16))
    xv, tv = batch(256)
    first = net.train_step(xv, tv, 0.0, want_loss=True)
    for _ in range(400):
        xb, tb = batch(256)
        net.train_step(xb, tb, 2e-3)
    last = net.train_step(xv, tv, 0.0, want_loss=True)
    print(f"loss {first:.4e} -> {last:.4e}")
    assert last < 0.05 * first
    p = net.get_params(0)
    net2 = _mlp(hip_lib, shape=shape, max_batch=256, seed=10)
    net2.set_params(p)
    np.testing.assert_array_equal(net2.get_params(0), p)
    np.testing.assert_array_equal(net2.get_params(1), net.get_params(1))
    y1, y2 = net.forward(xv), net2.forward(xv)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    net.close()
    net2.close()


def test_graph_fit_equals_eager_steps(hip_lib):
    """st_amd_mlp_fit replays one captured training step from a HIP graph;
    its parameters and loss must equal the same steps taken one eager
    train_step call at a time, bit for bit, after every call (each call
    captures a new graph and replays it; eager work runs in between)."""
    import torch
    shape = (8, 128, 3, 256)
    rng = np.random.default_rng(4)
    rows = 1000
    X = torch.from_numpy(rng.uniform(-1, 1, (rows, 8)).astype(np.float32)).cuda()
    tb = mo.bf16_bits(rng.normal(0, 0.5, (rows, 256)).astype(np.float32))
    T = torch.from_numpy(tb.view(np.int16)).cuda().view(torch.bfloat16)
    g = _mlp(hip_lib, shape, max_batch=256, seed=5)
    e = _mlp(hip_lib, shape, max_batch=256, seed=5)
    for steps in (7, 1, 9):
        idx = torch.from_numpy(rng.integers(0, rows, (steps, 200)).astype(np.int32)).cuda()
        lr = torch.from_numpy(np.linspace(1e-3, 5e-4, steps).astype(np.float32)).cuda()
        loss_g = g.fit(X, T, idx, lr, want_loss=True)
        loss_e = None
        for k in range(steps):
            rows_k = idx[k].long()
            loss_e = e.train_step(X.index_select(0, rows_k).contiguous(), T.index_select(0, rows_k).contiguous(),
                                  float(lr[k].item()), want_loss=(k == steps - 1))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(g.get_params(0), e.get_params(0))
        np.testing.assert_array_equal(g.get_params(1), e.get_params(1))
        assert loss_g == loss_e
    g.close()
    e.close()
