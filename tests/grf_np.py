"""CPU restatement of the GPU Gaussian-random-field generator (gft_init_grf),
test infrastructure: Philox4x32-10 written out with numpy integer arithmetic
(pinned by the published known-answer vectors in tests/test_grf.py),
Box-Muller, numpy FFTs, the same filter and normalisations."""
from __future__ import annotations

import numpy as np

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised over counter arrays; returns four uint64 arrays of 32-bit words."""
    c = [np.asarray(x, dtype=np.uint64) & MASK for x in (c0, c1, c2, c3)]
    k = [np.uint64(k0) & MASK, np.uint64(k1) & MASK]
    for r in range(10):
        if r:
            k = [(k[0] + np.uint64(W0)) & MASK, (k[1] + np.uint64(W1)) & MASK]
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k[0], p1 & MASK, (p0 >> np.uint64(32)) ^ c[3] ^ k[1], p0 & MASK]
    return c


def normal(seed, n, comp):
    """Standard normals of sites 0..n-1 of one component."""
    s = np.arange(n, dtype=np.uint64)
    o = philox4x32_10(s, np.full(n, comp, np.uint64), 0, 0, seed & 0xFFFFFFFF, seed >> 32)
    a = (o[1] << np.uint64(32)) | o[0]
    b = (o[3] << np.uint64(32)) | o[2]
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = ((b >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def component(n, d, seed, comp, slope, kmax):
    w = normal(seed, int(np.prod(n)), comp).reshape(n)
    wk = np.fft.fftn(w)
    k2 = 0.0
    for a in range(3):
        k = 2 * np.pi * np.fft.fftfreq(n[a], d=d[a])
        k2 = k2 + (k ** 2).reshape([-1 if b == a else 1 for b in range(3)])
    kk = np.sqrt(k2)
    amp = np.where((kk > 0) & (kk <= kmax), kk ** (slope / 2.0), 0.0)
    f = np.fft.ifftn(wk * amp).real
    f = f - f.mean()
    return (f / np.sqrt((f ** 2).mean())).ravel()


def gaussian_state(n, seed, sigma=0.05, vevs=(1.0,), ndoublets=1):
    """Two-slot arrays (FP64) of the state gft_init_gaussian builds: the vev
    plus sigma times the white noise of (site, component), unfiltered."""
    N = int(np.prod(n))
    s = np.zeros((2, N, 4 * ndoublets))
    for k in range(ndoublets):
        for m in range(4):
            s[1, :, 4 * k + m] = sigma * normal(seed, N, 4 * k + m) + (vevs[k] if m == 2 else 0.0)
    s[0] = s[1]
    v = np.zeros((2, N, 15))
    v[:, :, 1::5] = 1.0
    return s, v


def grf_state(n, d, seed, rms=0.1, slope=-1.0, kmax=None, vevs=(1.0,), ndoublets=1):
    """Two-slot arrays (FP64) of the state gft_init_grf builds."""
    kmax = np.pi / (2 * d[0]) if not kmax or kmax <= 0 else kmax
    N = int(np.prod(n))
    s = np.zeros((2, N, 4 * ndoublets))
    for k in range(ndoublets):
        comps = np.stack([component(n, d, seed, 4 * k + m, slope, kmax) for m in range(4)], -1)
        comps *= rms / np.sqrt((comps ** 2).sum(-1).mean())
        comps[:, 2] += vevs[k]
        s[1, :, 4 * k:4 * k + 4] = comps
    s[0] = s[1]
    v = np.zeros((2, N, 15))
    v[:, :, 1::5] = 1.0
    return s, v
