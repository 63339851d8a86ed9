"""GPU checks of the surrogate network's hand-written MFMA kernels (gft_mlp_*)
against the FP64 CPU reference of tests/mlp_np.py, which rounds to BF16 at the
same points as the GPU (inputs, weights, activations, dY, dZ).  The only
difference left is the GPU's FP32 accumulation (and the rare BF16 rounding
flip it causes), so forward outputs and gradients must agree closely; any
mistake in the MFMA lane layout, a transposed operand or a dropped term shows
up as an O(1) difference."""
from __future__ import annotations

import numpy as np
import pytest

from tests import mlp_np as M

pytestmark = pytest.mark.gpu

CFG = dict(n_in=8, width=256, n_hidden=4, n_heads=4, n_out=256)


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _net_and_params(seed=3, scale=1.0):
    net = M.MlpRef(**CFG)
    p = net.init(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    # Non-zero biases so every term of the layer is exercised.
    o = 0
    for sh in net.shapes:
        n = int(np.prod(sh))
        if len(sh) == 1:
            p[o:o + n] = 0.1 * rng.standard_normal(n)
        o += n
    return net, (p * scale).astype(np.float32)


def test_mlp_forward_matches_bf16_reference(rt):
    net, p = _net_and_params()
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=(37, 8)).astype(np.float32)   # not a multiple of 16: padded rows
    with rt.Mlp(**CFG, max_batch=32) as mlp:                   # two chunks
        mlp.set_params(p)
        assert np.array_equal(mlp.get_params(), p)
        y = mlp.forward(x)
    yr, _ = net.forward(p.astype(np.float64), x.astype(np.float64), bf=True)
    err = np.abs(y - yr).max() / np.abs(yr).max()
    print(f"forward: max |GPU - CPU| / max |y| = {err:.2e}")
    assert err < 1e-5


def test_mlp_gradients_match_bf16_reference(rt):
    net, p = _net_and_params(seed=4)
    rng = np.random.default_rng(6)
    B = 64
    x = rng.uniform(-1, 1, size=(B, 8)).astype(np.float32)
    t = rng.standard_normal((B, 256)).astype(np.float32)
    w = rng.poisson(1.0, size=(B, 4)).astype(np.float32)
    with rt.Mlp(**CFG, max_batch=64) as mlp:
        mlp.set_params(p)
        loss, g = mlp.gradients(x, t, w)
    lr, gr = net.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), w.astype(np.float64),
                           bf=True)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    o = 0
    for (name, sh) in zip([f"W{l}" if i % 2 == 0 else f"b{l}" for l in range(4) for i in range(2)] + ["Wh", "bh"],
                          net.shapes):
        n = int(np.prod(sh))
        r = np.linalg.norm(g[o:o + n] - gr[o:o + n]) / max(np.linalg.norm(gr[o:o + n]), 1e-30)
        print(f"  {name}: relative gradient difference {r:.2e}")
        assert r < 1e-3
        o += n
    print(f"loss {loss:.6e} vs {lr:.6e}; gradient relative difference {rel:.2e}")
    assert abs(loss - lr) <= 1e-5 * abs(lr)
    assert rel < 5e-4


def test_mlp_training_reduces_loss_and_is_deterministic(rt):
    """A smooth map R^8 -> R^256 (sums of sines of the inputs) is learnt; two
    runs with the same seed give bitwise the same parameters."""
    net = M.MlpRef(**CFG)
    p0 = net.init(7)
    rng = np.random.default_rng(8)
    n = 1024
    x = rng.uniform(-1, 1, size=(n, 8)).astype(np.float32)
    k = rng.standard_normal((8, 256))
    t = np.sin(x @ k * 0.5).astype(np.float32)
    w = np.ones((n, 4), np.float32)
    out = []
    for _ in range(2):
        with rt.Mlp(**CFG, max_batch=256) as mlp:
            mlp.set_params(p0)
            hist = mlp.train(x, t, w, batch=128, epochs=30, lr=2e-3, seed=11)
            out.append((hist, mlp.get_params()))
    hist = out[0][0]
    print(f"loss per epoch: {hist[0]:.4e} -> {hist[9]:.4e} -> {hist[-1]:.4e}")
    assert hist[-1] < 0.1 * hist[0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][0], out[1][0])
