"""CPU checks of the FP64 reference MLP that the GPU surrogate network is
compared against (tests/mlp_np.py): its analytic gradient against central
differences, closed-form cases, and the BF16 rounding helper."""
from __future__ import annotations

import numpy as np

from tests import mlp_np as M


def test_bf16_rounding_is_nearest_even():
    # BF16 keeps 8 significant bits: the spacing at 1 is 2^-7.
    x = np.array([1.0, 1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -9, -2.5, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 1.0, 1.0 + 2.0 ** -6])   # ties go to the even neighbour
    assert np.array_equal(M.bf16(x), want)
    assert abs(M.bf16(np.array([3.0e-30]))[0] - 3.0e-30) <= 3.0e-30 * 2.0 ** -8


def test_gelu_closed_forms():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    z = np.array([-30.0, 30.0])
    assert np.allclose(M.gelu(z), [0.0, 30.0]) and np.allclose(M.gelu_grad(z), [0.0, 1.0])
    # derivative by central differences
    z = np.linspace(-4, 4, 17)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(z)).max() < 1e-8


def test_reference_gradient_matches_central_differences():
    net = M.MlpRef(n_in=3, width=8, n_hidden=2, n_heads=2, n_out=5)
    rng = np.random.default_rng(1)
    p = rng.standard_normal(net.n_params) * 0.5
    x = rng.standard_normal((6, 3))
    t = rng.standard_normal((6, 5))
    w = rng.integers(0, 3, size=(6, 2)).astype(float)
    loss, g = net.loss_grad(p, x, t, w)
    h = 1e-6
    for i in rng.choice(net.n_params, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (net.loss_grad(pp, x, t, w)[0] - net.loss_grad(pm, x, t, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_closed_form_single_layer():
    """W0 = 0 makes every hidden unit GELU(b0); the heads then output
    Wh . GELU(b0) + bh for every input."""
    net = M.MlpRef(n_in=2, width=4, n_hidden=1, n_heads=1, n_out=3)
    b0 = np.array([0.5, -1.0, 2.0, 0.0])
    Wh = np.arange(12, dtype=float).reshape(3, 4) / 10
    bh = np.array([1.0, -1.0, 0.5])
    p = np.concatenate([np.zeros(8), b0, Wh.ravel(), bh])
    y, _ = net.forward(p, np.array([[3.0, -7.0], [0.1, 0.2]]))
    want = Wh @ M.gelu(b0) + bh
    assert np.allclose(y[0, 0], want) and np.allclose(y[1, 0], want)
    # Zero loss at t = y, and zero gradient.
    loss, g = net.loss_grad(p, np.array([[3.0, -7.0]]), y[:1, 0], np.ones((1, 1)))
    assert loss == 0.0 and np.abs(g).max() == 0.0
