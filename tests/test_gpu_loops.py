"""Pins of the Wilson-loop observables (gft_wilson_loops; no reference
counterpart -- the reference measures the 1x1 loop only through its magnetic
energy, WilsonLoops/Source/StandardModel.cpp:397-401).

  * Closed forms on abelian configurations: a constant field strength theta in
    every plane, from the U(1) phases y_j(x) = theta * sum_{i<j} x_i, gives
    W11_u1 = cos(theta) and W22_u1 = cos(4 theta); the same pattern of angles on
    SU(2) links exp(i alpha n.sigma) with a fixed axis n (commuting links) gives
    W11_su2 = cos(alpha), W22_su2 = cos(4 alpha).  A loop with a wrong side
    length, a dropped link or a wrong sign on one phase gives another cosine.
  * Gauge invariance: a random site-wise transformation G(x) = W(x) e^{i a(x)}
    (far from the identity) leaves every loop unchanged; a loop whose links are
    multiplied in a wrong order or with a wrong conjugation is not invariant
    for non-commuting links.
  * An FP64 restatement of the path products on a hot random state.
  * The 2x2 U(1) loop is 1 on identity links and cos of the summed phases of
    its 8 links otherwise.
"""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, make_desc
from tests import lattice_np as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _state(n, y, q, rng=None):
    """Two-slot arrays with both slots equal (E = 0, P = 0): links y[N,3], q[N,3,4]."""
    N = int(np.prod(n))
    s = np.zeros((2, N, 4))
    s[:, :, 2] = 1.0
    if rng is not None:
        s[:] = s[1] + rng.normal(size=(N, 4)) * 0.1
    v = np.zeros((2, N, 3, 5))
    v[:, :, :, 0] = y
    v[:, :, :, 1:] = q
    return s, v.reshape(2, N, 15)


def _loops(rt, form, n, s, v):
    with rt.Evolver(make_desc(form=form, n=n)) as ev:
        ev.import_two_slot(s, v, 0)
        w = ev.wilson_loops()
    return np.array([w.w11_su2, w.w11_u1, w.w22_su2, w.w22_u1])


def _coords(n):
    x = np.indices(n).reshape(3, -1).T   # site s = (x * ny + y) * nz + z
    return x


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 2e-14), (GFT_FORM_MOMENTUM_F32, 1e-6)],
                         ids=["f64", "f32"])
@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.48, -0.6, 0.64)], ids=["sigma3", "tilted"])
def test_abelian_constant_field_closed_forms(rt, form, tol, axis):
    n = (12, 12, 10)
    x = _coords(n)
    theta, alpha = 2 * np.pi * 5 / 12, 2 * np.pi / 12      # theta * n_i and alpha * n_i are multiples of 2 pi
    N = x.shape[0]
    y = np.zeros((N, 3))
    ang = np.zeros((N, 3))
    for j in range(3):
        lin = x[:, :j].sum(axis=1) if j else np.zeros(N)
        y[:, j] = theta * lin
        ang[:, j] = alpha * lin
    nax = np.array(axis) / np.linalg.norm(axis)
    q = np.concatenate([np.cos(ang)[..., None], np.sin(ang)[..., None] * nax], axis=-1)
    s, v = _state(n, y, q)
    w = _loops(rt, form, n, s, v)
    expect = np.array([np.cos(alpha), np.cos(theta), np.cos(4 * alpha), np.cos(4 * theta)])
    print(f"form {form}, axis {axis}: loops {w} expected {expect}")
    np.testing.assert_allclose(w, expect, rtol=0, atol=tol)


def _hot_state(n, seed):
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    y = rng.uniform(-1.0, 1.0, size=(N, 3))
    q = L.random_su2(rng, (N, 3), 0.7)
    return _state(n, y, q, rng)


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-13), (GFT_FORM_MOMENTUM_F32, 2e-6)],
                         ids=["f64", "f32"])
def test_loops_gauge_invariant(rt, form, tol):
    n = (8, 6, 10)
    s, v = _hot_state(n, 11)
    rng = np.random.default_rng(12)
    N = int(np.prod(n))
    a = rng.uniform(-np.pi, np.pi, size=N)
    g = L.random_su2(rng, (N,), 3.0)           # far from the identity
    s2, v2 = L.gauge_transform(s, v, n, a, g)
    w1 = _loops(rt, form, n, s, v)
    w2 = _loops(rt, form, n, s2, v2)
    print(f"form {form}: loops {w1} after a random gauge transformation {w2} (|diff| {np.abs(w1 - w2).max():.1e})")
    assert np.abs(w1 - 1).min() > 0.05           # hot: far from the trivial value
    np.testing.assert_allclose(w2, w1, rtol=0, atol=tol)
    # A product in the wrong order is not invariant: the same links taken as
    # U_i(x) U_j(x+i) U_i(x+j)^dag U_j(x)^dag with the other convention move.
    bad1 = _np_loops(n, v, reverse=True)
    bad2 = _np_loops(n, v2, reverse=True)
    assert np.abs(bad1[[0, 2]] - bad2[[0, 2]]).max() > 1e-3


def _np_loops(n, v, reverse=False):
    """FP64 restatement: mean over sites and the three planes (i<j) of
    1/2 tr and cos of the LxL loop x -> x+Li -> x+Li+Lj -> x+Lj -> x, links
    multiplied in path order (the product convention of lattice_np.qmul, in
    which U_i(x) carries the fibre at x to x+i)."""
    N = int(np.prod(n))
    vv = v[1].reshape(N, 3, 5)
    y, q = vv[..., 0], vv[..., 1:]
    out = np.zeros(4)
    for Lsz in (1, 2):
        su2, u1 = 0.0, 0.0
        for i in range(3):
            for j in range(i + 1, 3):
                W = np.zeros((N, 4))
                W[:, 0] = 1.0
                ph = np.zeros(N)
                # steps: (direction, sign) along the path; position tracked as a shift
                pos = [0, 0, 0]
                path = [(i, +1)] * Lsz + [(j, +1)] * Lsz + [(i, -1)] * Lsz + [(j, -1)] * Lsz
                for d, sg in path:
                    if sg < 0:
                        pos[d] -= 1
                    qq = q[:, d]
                    yy = y[:, d]
                    for ax in range(3):
                        if pos[ax]:
                            qq = L.shift(qq, n, ax, pos[ax])
                            yy = L.shift(yy, n, ax, pos[ax])
                    if sg > 0:
                        W = L.qmul(W, qq) if reverse else L.qmul(qq, W)
                        ph += yy
                        pos[d] += 1
                    else:
                        W = L.qmul(W, L.qconj(qq)) if reverse else L.qmul(L.qconj(qq), W)
                        ph -= yy
                nrm = np.linalg.norm(W, axis=-1)
                su2 += (W[:, 0] / nrm).sum()
                u1 += np.cos(ph).sum()
        out[2 * (Lsz - 1)] = su2 / (3 * N)
        out[2 * (Lsz - 1) + 1] = u1 / (3 * N)
    return np.array([out[0], out[1], out[2], out[3]])


def test_loops_match_fp64_restatement(rt):
    n = (8, 6, 10)
    s, v = _hot_state(n, 21)
    w = _loops(rt, GFT_FORM_MOMENTUM_F64, n, s, v)
    ref = _np_loops(n, v)
    print(f"GPU {w}, restatement {ref}")
    np.testing.assert_allclose(w, ref, rtol=0, atol=1e-13)


def test_u1_2x2_loop_identity_and_single_link(rt):
    """Identity links: every loop is 1.  One U(1) phase y on link (x0, i = 0):
    it belongs to 2 of the 1x1 and 4 of the 2x2 loops in each of its 2 planes
    ((0,1) and (0,2)), so mean cos = 1 - k (1 - cos y) / (3 N) with k = 4 (1x1)
    and k = 8 (2x2)."""
    n = (6, 6, 6)
    N = 216
    y = np.zeros((N, 3))
    q = np.zeros((N, 3, 4))
    q[..., 0] = 1.0
    s, v = _state(n, y, q)
    w = _loops(rt, GFT_FORM_MOMENTUM_F64, n, s, v)
    np.testing.assert_allclose(w, 1.0, rtol=0, atol=1e-15)
    y[37, 0] = 0.9
    s, v = _state(n, y, q)
    w = _loops(rt, GFT_FORM_MOMENTUM_F64, n, s, v)
    c = 1 - np.cos(0.9)
    np.testing.assert_allclose(w[1], 1 - 4 * c / (3 * N), rtol=0, atol=1e-15)
    np.testing.assert_allclose(w[3], 1 - 8 * c / (3 * N), rtol=0, atol=1e-15)
    np.testing.assert_allclose(w[[0, 2]], 1.0, rtol=0, atol=1e-15)
