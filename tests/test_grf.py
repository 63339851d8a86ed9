"""CPU checks of the Philox4x32-10 restatement and of the GPU-generator's CPU
restatement (tests/grf_np.py)."""
from __future__ import annotations

import numpy as np

from tests import grf_np as G
from tests.test_ic import _spectrum


def test_philox_known_answers():
    """The published Philox4x32-10 known-answer vectors (Random123)."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
             ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
             ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
              (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in cases:
        got = G.philox4x32_10(*[np.array([c]) for c in ctr], *key)
        assert tuple(int(x[0]) for x in got) == want


def test_philox_normals_are_standard():
    z = G.normal(12345, 1 << 18, 3)
    assert abs(z.mean()) < 4 / np.sqrt(len(z)) and abs(z.std() - 1) < 4 / np.sqrt(2 * len(z))
    assert abs((z ** 4).mean() - 3) < 0.05
    z2 = G.normal(12345, 1 << 18, 4)
    assert abs(np.corrcoef(z, z2)[0, 1]) < 4 / np.sqrt(len(z))


def test_restatement_field_statistics():
    n, d = (48, 32, 24), (0.5, 0.5, 0.5)
    s, v = G.grf_state(n, d, seed=7, rms=0.1, slope=-1.0)
    dphi = s[1] - np.array([0.0, 0.0, 1.0, 0.0])
    assert abs(np.sqrt((dphi ** 2).sum(axis=1).mean()) - 0.1) < 1e-12
    f = G.component(n, d, 7, 0, -1.0, np.pi)
    k, p = _spectrum(f.reshape(n), d)
    assert p[k > np.pi * (1 + 1e-9)].sum() <= 1e-20 * p.sum()
