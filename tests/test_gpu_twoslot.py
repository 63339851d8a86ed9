"""GPU parity of the two-slot (reference layout) evolver against the CPU oracle.

Tolerances:
  * FP64 storage: <= 1e-12 relative (field level) after 100 steps of the 32^3
    SM workload -- the only differences are the device's sin/cos/asin.
  * FP32 storage (the reference's own): after a few steps at most a few float
    ulps differ, from 1-ulp differences of the FP64 sin/cos/asin before the
    float rounding.
  * Diagnostics (FP64 compensated sums): <= 1e-10 relative, with an absolute
    floor of 1e-11 x the total energy per column: the magnetic and electric
    densities are the reference's 1 - cos(theta) and 1 - P0 (WilsonLoops/Source/
    StandardModel.cpp:397-401, 440-446), which cancel for small angles and turn
    a 1-ulp difference of cos into ~1e-9 relative of that column.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64, GFT_MODEL_2HDM, GFT_MODEL_SM,
                                 TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def gaussian_ic(n, Ns=4, sigma=0.05, seed=20261203):
    """Config of the 32^3 workload: phi at the vev (0,0,1,0) plus iid Gaussian
    noise of sigma per real component, both slots equal, links identity."""
    rng = np.random.default_rng(seed)
    N = int(np.prod(n))
    s = np.zeros((2, N, Ns))
    s[1] = rng.normal(size=(N, Ns)) * sigma
    for k in range(Ns // 4):
        s[1][:, 4 * k + 2] += 1.0
    s[0] = s[1]
    return s, O.identity_links(N, np.float64)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_twoslot_f64_matches_oracle_32cube_100_steps(rt):
    n = (32, 32, 32)
    d = make_desc(form=GFT_FORM_TWO_SLOT_F64, n=n)
    s, v = gaussian_ic(n)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(100, energy_freq=10, gauss_freq=10)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 100
    _, orows, _ = O.evolve(d, s, v, 100, energy_freq=10, gauss_freq=10, nthreads=8)
    rs, rv = _rel(gs, s), _rel(gv, v)
    print(f"FP64 two-slot vs oracle after 100 steps: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= 1e-12 and rv <= 1e-12
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter
        np.testing.assert_allclose(np.array(a.energy[:]), np.array(b.energy[:]), rtol=1e-10,
                                   atol=1e-11 * abs(b.energy[0]))
        np.testing.assert_allclose(np.array(a.gauss_scale[:]), np.array(b.gauss_scale[:]), rtol=1e-10)
        # |G| itself is at rounding level (~1e-15 x scale); compare on that scale.
        assert np.abs(np.array(a.gauss_int[:]) - np.array(b.gauss_int[:])).max() <= 1e-11 * max(b.gauss_scale[:])


@pytest.mark.parametrize("model,tag,pot2", [(GFT_MODEL_SM, "sm", None), (GFT_MODEL_2HDM, "2hdm", TWO_HDM_POTENTIAL)],
                         ids=["sm", "2hdm"])
def test_twoslot_f32_matches_reference_golden(rt, model, tag, pot2):
    """FP32 storage from the reference harness's initial state; the golden
    fields after steps 1, 3 and 6 were written by the reference's own code."""
    g = np.load(os.path.join(GOLD, f"harness_{tag}.npz"))
    n = (4, 5, 6)
    d = make_desc(model=model, form=GFT_FORM_TWO_SLOT_F32, n=n, pot2=pot2)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(g["scalar0"], g["vec0"], 0)
        for k in range(6):
            ev.evolve(1, energy_freq=3, gauss_freq=3)
            if f"scalar_{k}" in g:
                s, v, _ = ev.export_two_slot(np.float32)
                ds = np.abs(s - g[f"scalar_{k}"]).max()
                dv = np.abs(v - g[f"vec_{k}"]).max()
                print(f"{tag} step {k}: max |GPU - reference| scalar {ds:.2e} links {dv:.2e}")
                # A few float ulps at |field| <= ~2.
                assert ds <= 2e-6 and dv <= 2e-6


def test_twoslot_2hdm_f64_matches_oracle(rt):
    n = (8, 6, 10)
    d = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_TWO_SLOT_F64, n=n, pot2=TWO_HDM_POTENTIAL)
    s, v = L.near_vacuum_state(n, 8, seed=9, amp=0.1, mom=0.01, link_eps=0.05, vev=0.7)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(40, energy_freq=8, gauss_freq=8)
        gs, gv, _ = ev.export_two_slot(np.float64)
    _, orows, _ = O.evolve(d, s, v, 40, energy_freq=8, gauss_freq=8, nthreads=8)
    assert _rel(gs, s) <= 1e-12 and _rel(gv, v) <= 1e-12
    for a, b in zip(rows, orows):
        np.testing.assert_allclose(np.array(a.energy[:]), np.array(b.energy[:]), rtol=1e-10,
                                   atol=1e-11 * abs(b.energy[0]))


def test_twoslot_damping_and_ragged_lattice(rt):
    """Damped steps (Model::update) on a lattice with unequal sides and spacings."""
    n = (5, 7, 3)
    d = make_desc(form=GFT_FORM_TWO_SLOT_F64, n=n, d=(0.5, 0.4, 0.6), damping=1.0, nt_damped=5)
    s, v = L.near_vacuum_state(n, 4, seed=2, amp=0.2, mom=0.02, link_eps=0.1)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 3)
        ev.evolve(6)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 9
    O.evolve(d, s, v, 6, time_iter=3)
    assert _rel(gs, s) <= 1e-13 and _rel(gv, v) <= 1e-13


def test_twoslot_domain_error_is_reported(rt):
    """|R| > 1 (a huge time step) is reported as GFT_EDOMAIN, not silently NaN."""
    n = (4, 4, 4)
    d = make_desc(form=GFT_FORM_TWO_SLOT_F64, n=n, dt=3.0)
    s, v = L.near_vacuum_state(n, 4, seed=2, amp=0.5, mom=0.2, link_eps=0.5)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        with pytest.raises(rt.GftError) as e:
            ev.evolve(3)
        assert e.value.code == -4
