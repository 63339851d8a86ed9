"""GPU parity of the global O(N) model: N real scalars with the plain
Laplacian and an SO_N (or Double_SO_N) potential, no gauge links
(Gradients/Source/Global.cpp with the null Wilson loop, WilsonLoops/Source/
WilsonLoop.cpp:22-85).  It runs in the two-slot forms:

  * two-slot FP32 against the fields the reference's own code wrote from the
    same state (harness_global_n<N>.npz, unequal spacings): a few float ulps;
  * two-slot FP64 against the FP64 oracle: <= 1e-12 relative (fields),
    energies <= 1e-10 relative;
  * no Gauss law: measuring rows carry energies only.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gftevolver_amd._abi import GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64, TWO_HDM_POTENTIAL
from tests import oracle_lib as O
from tests.test_oracle import GLOBAL, global_desc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


@pytest.mark.parametrize("ns,pot2", GLOBAL, ids=["n3-son", "n8-double-son"])
def test_twoslot_f32_global_model_matches_reference_golden(rt, ns, pot2):
    g = np.load(os.path.join(GOLD, f"harness_global_n{ns}.npz"))
    d = global_desc((4, 5, 6), ns, pot2, d=(0.5, 0.4, 0.6), form=GFT_FORM_TWO_SLOT_F32)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(g["scalar0"], np.zeros(0, np.float32), 0)
        for k in range(6):
            ev.evolve(1)
            if f"scalar_{k}" in g:
                s, v, _ = ev.export_two_slot(np.float32)
                assert v.size == 0
                ds = np.abs(s - g[f"scalar_{k}"]).max()
                print(f"N={ns} step {k}: max |GPU - reference| {ds:.2e}")
                assert ds <= 2e-6


@pytest.mark.parametrize("ns,pot2", GLOBAL + [(1, None), (6, None)], ids=["n3-son", "n8-double-son", "n1", "n6"])
def test_twoslot_f64_global_model_matches_oracle(rt, ns, pot2):
    n = (10, 8, 12)
    d = global_desc(n, ns, pot2, d=(0.5, 0.45, 0.55), form=GFT_FORM_TWO_SLOT_F64)
    N = int(np.prod(n))
    rng = np.random.default_rng(40 + ns)
    s = np.zeros((2, N, ns))
    s[1] = rng.normal(size=(N, ns)) * 0.3
    s[1][:, 0] += 1.0
    s[0] = s[1] + rng.normal(size=(N, ns)) * 0.02
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, np.zeros(0), 0)
        rows = ev.evolve(40, energy_freq=10, gauss_freq=10)
        gs, _, it = ev.export_two_slot(np.float64)
    assert it == 40
    _, orows, _ = O.evolve(d, s, None, 40, energy_freq=10, nthreads=8)
    rs = np.abs(gs - s).max() / np.abs(s).max()
    print(f"N={ns}: FP64 two-slot vs oracle after 40 steps: {rs:.2e}")
    assert rs <= 1e-12
    assert len(rows) == len(orows) == 4
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter and a.has_energy and not a.has_gauss
        np.testing.assert_allclose(np.array(a.energy[:]), np.array(b.energy[:]), rtol=1e-10,
                                   atol=1e-12 * abs(b.energy[0]))
        assert a.energy[4] == a.energy[5] == 0.0
