"""CPU checks of the initial-state generators (gftevolver_amd/ic.py), which
have no reference counterpart: the statistics fix them.  A Gaussian random
field filtered to P(k) ~ |k|^slope for 0 < |k| <= k_max must have zero mean,
the requested rms, no power above k_max, the requested spectral slope, and
Gaussian one-point statistics."""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd import ic


def _spectrum(f, d):
    n = f.shape
    fk = np.abs(np.fft.fftn(f)) ** 2
    k2 = sum((2 * np.pi * np.fft.fftfreq(n[a], d=d[a])).reshape([-1 if b == a else 1 for b in range(3)]) ** 2
             for a in range(3))
    return np.sqrt(k2).ravel(), fk.ravel()


@pytest.mark.parametrize("slope", [-1.0, -2.0, 0.0])
def test_grf_spectrum_slope_cutoff_and_moments(slope):
    n, d = (64, 64, 64), (0.5, 0.5, 0.5)
    kmax = np.pi / (2 * d[0])
    f = ic.grf_component(n, d, np.random.default_rng(11), slope=slope)
    assert abs(f.mean()) < 1e-12 and abs(np.sqrt((f ** 2).mean()) - 1.0) < 1e-12
    k, p = _spectrum(f, d)
    tot = p.sum()
    assert p[k > kmax * (1 + 1e-9)].sum() <= 1e-20 * tot       # nothing above the cut-off
    assert p[k == 0].sum() <= 1e-20 * tot                        # zero mean
    # Binned spectrum: the slope of log P against log k over the band.
    edges = np.geomspace(k[k > 0].min() * 1.01, kmax, 12)
    kc, pc = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        m = (k >= a) & (k < b)
        if m.sum() > 50:
            kc.append(np.log(k[m]).mean())
            pc.append(np.log(p[m].mean()))
    fit = np.polyfit(kc, pc, 1)[0]
    print(f"slope {slope}: fitted {fit:.3f} over {len(kc)} bins")
    assert abs(fit - slope) < 0.1
    # One-point statistics of a Gaussian field.
    z = f.ravel()
    assert abs((z ** 3).mean()) < 0.05 and abs((z ** 4).mean() - 3.0) < 0.1


def test_grf_state_layout_amplitude_and_seed():
    n = (16, 12, 20)
    s, v = ic.grf_state(n, rms=0.1, vev=1.0, seed=5, dtype=np.float64)
    N = 16 * 12 * 20
    assert s.shape == (2, N, 4) and v.shape == (2, N, 15)
    assert np.array_equal(s[0], s[1])
    dphi = s[1] - np.array([0.0, 0.0, 1.0, 0.0])
    assert abs(np.sqrt((dphi ** 2).sum(axis=1).mean()) - 0.1) < 1e-12   # rms over all four components
    assert np.abs(dphi.mean(axis=0)).max() < 1e-12
    # identity links, zero phases, both slots
    assert np.array_equal(v[..., 1::5], np.ones((2, N, 3))) and not v[..., 0::5].any()
    assert not v[..., 2::5].any() and not v[..., 3::5].any() and not v[..., 4::5].any()
    s2, _ = ic.grf_state(n, rms=0.1, vev=1.0, seed=5, dtype=np.float64)
    s3, _ = ic.grf_state(n, rms=0.1, vev=1.0, seed=6, dtype=np.float64)
    assert np.array_equal(s, s2) and not np.array_equal(s, s3)


def test_grf_state_two_doublets():
    s, _ = ic.grf_state((8, 8, 8), rms=0.1, ndoublets=2, vevs=(0.8, 0.6), seed=3, dtype=np.float64)
    assert s.shape == (2, 512, 8)
    for k, vev in enumerate((0.8, 0.6)):
        dphi = s[1][:, 4 * k:4 * k + 4] - np.array([0.0, 0.0, vev, 0.0])
        assert abs(np.sqrt((dphi ** 2).sum(axis=1).mean()) - 0.1) < 1e-12


def test_gaussian_state_statistics():
    s, _ = ic.gaussian_state((32, 32, 32), sigma=0.05, vev=1.0, seed=1)
    d = s[1] - np.array([0.0, 0.0, 1.0, 0.0])
    assert np.abs(d.mean(axis=0)).max() < 0.05 * 4 / np.sqrt(32 ** 3)
    assert np.abs(d.std(axis=0) - 0.05).max() < 0.05 * 4 / np.sqrt(2 * 32 ** 3)


def test_ensemble_fixed_tiles_lookup():
    """Fixed ensemble tiles are looked up by member shape or by a cubic
    member's side, and a size without a tile is refused before any launch."""
    from gftevolver_amd.ensemble import fixed_tiles
    sizes = [(32, 32, 32), (48, 48, 48), (16, 24, 32)]
    got = fixed_tiles(sizes, {32: (32, 8), 48: [16, 16], (16, 24, 32): (16, 16)})
    assert got == {(32, 32, 32): (32, 8), (48, 48, 48): (16, 16), (16, 24, 32): (16, 16)}
    with pytest.raises(ValueError, match="no tile given"):
        fixed_tiles(sizes, {32: (32, 8), 48: (16, 16), 16: (16, 16)})   # a side does not name a non-cubic shape
