"""Small numpy helpers for tests: quaternion algebra, gauge transformations and
seeded states in the reference two-slot layout (scalar[2][N][Ns], vec[2][N][15]).

Conventions (Gradients/Source/StandardModel.cpp:72-100, WilsonLoops/Source/
StandardModel.cpp:76-191): a link U_i(x) carries the fibre at x to x+i, so under
a gauge transformation G(x) = W(x) e^{i a(x)} (W in SU(2) as a unit quaternion,
a the U(1) phase)
    phi(x)   -> W(x) e^{i a(x)} phi(x)
    y_i(x)   -> y_i(x) + a(x+i) - a(x)
    q_i(x)   -> W(x+i) q_i(x) W(x)^dagger
"""
from __future__ import annotations

import numpy as np


def qmul(a, b):
    """Quaternion product in the reference's SU2Product convention
    (U = c0 + i c.sigma; WilsonLoops/Source/StandardModel.cpp:165-176)."""
    a0, a1, a2, a3 = np.moveaxis(a, -1, 0)
    b0, b1, b2, b3 = np.moveaxis(b, -1, 0)
    return np.stack([
        a0 * b0 - (a1 * b1 + a2 * b2 + a3 * b3),
        a0 * b1 + a1 * b0 - a2 * b3 + a3 * b2,
        a0 * b2 + a2 * b0 - a3 * b1 + a1 * b3,
        a0 * b3 + a3 * b0 - a1 * b2 + a2 * b1,
    ], axis=-1)


def qconj(a):
    out = a.copy()
    out[..., 1:] *= -1
    return out


def su2_act(q, h):
    """Real 4x4 action of U = c0 + i c.sigma on a doublet (f0+if1, f2+if3)
    written independently as complex 2x2 matrices (Pauli basis)."""
    c0, c1, c2, c3 = np.moveaxis(q, -1, 0)
    U = np.empty(q.shape[:-1] + (2, 2), complex)
    U[..., 0, 0] = c0 + 1j * c3
    U[..., 0, 1] = 1j * c1 + c2
    U[..., 1, 0] = 1j * c1 - c2
    U[..., 1, 1] = c0 - 1j * c3
    z = np.stack([h[..., 0] + 1j * h[..., 1], h[..., 2] + 1j * h[..., 3]], -1)
    w = np.einsum("...ij,...j->...i", U, z)
    return np.stack([w[..., 0].real, w[..., 0].imag, w[..., 1].real, w[..., 1].imag], -1)


def u1_act(a, h):
    z = np.stack([h[..., 0] + 1j * h[..., 1], h[..., 2] + 1j * h[..., 3]], -1) * np.exp(1j * a)[..., None]
    return np.stack([z[..., 0].real, z[..., 0].imag, z[..., 1].real, z[..., 1].imag], -1)


def shift(arr, n, axis, delta):
    """Value at x + delta*e_axis for per-site arrays of shape (N, ...)."""
    a = arr.reshape(tuple(n) + arr.shape[1:])
    return np.roll(a, -delta, axis=axis).reshape(arr.shape)


def random_su2(rng, shape, eps):
    q = rng.normal(size=shape + (4,)) * eps
    q[..., 0] += 1.0
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def gauge_transform(scalar, vec, n, a, w):
    """Apply G(x) = W(x) e^{i a(x)} to both slots of a two-slot state."""
    Ns = scalar.shape[-1]
    s = scalar.astype(np.float64).copy()
    v = vec.astype(np.float64).reshape(2, -1, 3, 5).copy()
    for t in range(2):
        for k in range(Ns // 4):
            s[t][:, 4 * k:4 * k + 4] = su2_act(w, u1_act(a, s[t][:, 4 * k:4 * k + 4]))
        for i in range(3):
            ap = shift(a, n, i, +1)
            wp = shift(w, n, i, +1)
            v[t][:, i, 0] = v[t][:, i, 0] + ap - a
            v[t][:, i, 1:] = qmul(qmul(wp, v[t][:, i, 1:]), qconj(w))
    return s, v.reshape(2, -1, 15)


def near_vacuum_state(n, Ns=4, seed=0, amp=0.1, mom=0.01, link_eps=0.05, dtype=np.float64, vev=1.0):
    """Doublet(s) at the vev (0, 0, vev, 0) plus Gaussian noise, links near identity,
    small random t-dt offsets."""
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    s = np.zeros((2, N, Ns))
    s[1] = rng.normal(size=(N, Ns)) * amp
    for k in range(Ns // 4):
        s[1][:, 4 * k + 2] += vev
    s[0] = s[1] + rng.normal(size=(N, Ns)) * mom
    v = np.zeros((2, N, 3, 5))
    v[1, ..., 0] = rng.normal(size=(N, 3)) * link_eps
    v[1, ..., 1:] = random_su2(rng, (N, 3), link_eps)
    v[0, ..., 0] = v[1, ..., 0] + rng.normal(size=(N, 3)) * mom
    v[0, ..., 1:] = qmul(random_su2(rng, (N, 3), mom), v[1, ..., 1:])
    return s.astype(dtype), v.reshape(2, N, 15).astype(dtype)
