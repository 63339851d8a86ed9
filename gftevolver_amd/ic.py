"""Initial states for the evolver, in the reference two-slot layout.

The reference has only RandomUniform and zero initial conditions
(InitialConditions/Source/RandomUniform.cpp:60-77); the workloads need
  * iid Gaussian fluctuations around the vev (32^3 workload),
  * Gaussian random fields with a power-law spectrum cut off at k_max
    (256^3, 2HDM, ensemble and surrogate workloads).
Every state has identity links, E = 0 and zero scalar momentum (both slots
equal), so the Gauss law holds exactly at t = 0.

Conventions: the vev sits in component 2, phi = (0, 0, v, 0) + dphi (the SO_N
potential is O(4) symmetric, so the direction is a choice).  "rms amplitude A"
means sqrt(<sum_c dphi_c^2>) = A over the lattice, all four components
together.
"""
from __future__ import annotations

import numpy as np

NV = 15


def identity_links(nsites: int, dtype=np.float32) -> np.ndarray:
    v = np.zeros((2, nsites, NV), dtype)
    v[:, :, 1::5] = 1.0
    return v


def grf_component(n, d, rng, slope=-1.0, kmax=None):
    """One real Gaussian random field on the periodic lattice n with spacing d,
    power spectrum P(k) ~ |k|^slope for 0 < |k| <= kmax (zero elsewhere, zero
    mean), unit rms.  White noise is filtered by sqrt(P) in Fourier space."""
    n = tuple(int(x) for x in n)
    kmax = np.pi / (2 * d[0]) if kmax is None else kmax
    w = rng.standard_normal(n)
    wk = np.fft.rfftn(w)
    k2 = np.zeros(wk.shape)
    for a in range(3):
        if a < 2:
            k = 2 * np.pi * np.fft.fftfreq(n[a], d=d[a])
        else:
            k = 2 * np.pi * np.fft.rfftfreq(n[a], d=d[a])
        shape = [1, 1, 1]
        shape[a] = -1
        k2 = k2 + (k ** 2).reshape(shape)
    kk = np.sqrt(k2)
    amp = np.zeros_like(kk)
    m = (kk > 0) & (kk <= kmax)
    amp[m] = kk[m] ** (slope / 2.0)
    f = np.fft.irfftn(wk * amp, s=n, axes=(0, 1, 2))
    f -= f.mean()
    return f / np.sqrt((f ** 2).mean())


def grf_state(n, d=(0.5, 0.5, 0.5), rms=0.1, vev=1.0, seed=20261203, ndoublets=1, vevs=None, slope=-1.0,
              kmax=None, dtype=np.float32):
    """Two-slot state: each doublet phi_k = (0, 0, vev_k, 0) + GRF fluctuations
    of total rms amplitude `rms`, both slots equal, identity links."""
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    Ns = 4 * ndoublets
    vevs = [vev] * ndoublets if vevs is None else list(vevs)
    s = np.zeros((2, N, Ns), dtype)
    for k in range(ndoublets):
        comps = np.stack([grf_component(n, d, rng, slope, kmax).reshape(-1) for _ in range(4)], -1)
        comps *= rms / np.sqrt((comps ** 2).sum(-1).mean())
        comps[:, 2] += vevs[k]
        s[1, :, 4 * k:4 * k + 4] = comps
    s[0] = s[1]
    return s, identity_links(N, dtype)


def gaussian_state(n, sigma=0.05, vev=1.0, seed=20261203, dtype=np.float64):
    """phi at the vev plus iid Gaussian noise of `sigma` per real component."""
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    s = np.zeros((2, N, 4), dtype)
    s[1] = rng.normal(size=(N, 4)) * sigma
    s[1][:, 2] += vev
    s[0] = s[1]
    return s, identity_links(N, dtype)
