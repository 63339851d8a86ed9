"""GPU parity of the Gaussian-random-field generator (gft_init_grf) with its
CPU restatement (tests/grf_np.py, whose Philox is pinned by the published
known-answer vectors): same white noise bit for bit (counter-based), FFTs
and normalisations to FP64 rounding."""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_FORM_TWO_SLOT_F32,
                                 GFT_FORM_TWO_SLOT_F64, GFT_MODEL_2HDM, GFT_MODEL_SM, TWO_HDM_POTENTIAL, make_desc)
from tests import grf_np as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


# FP32 momentum form: phi stored as its FP32 difference from the vev, at most
# ulp(|phi - vev|) / 2 <= 1.5e-8 for |phi - vev| < 0.5.
@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-12), (GFT_FORM_TWO_SLOT_F64, 1e-12),
                                      (GFT_FORM_MOMENTUM_F32, 1.5e-8), (GFT_FORM_TWO_SLOT_F32, 1e-7)],
                         ids=["mom64", "two64", "mom32", "two32"])
def test_grf_matches_restatement_sm(rt, form, tol):
    n, d = (24, 18, 32), (0.5, 0.4, 0.6)
    desc = make_desc(form=form, n=n, d=d)
    with rt.Evolver(desc) as ev:
        ev.init_grf(seed=0x1234_5678_9ABC, rms=0.1, slope=-1.5, kmax=0.0, vevs=(1.0, 1.0))
        gs, gv, it = ev.export_two_slot(np.float64)
        obs = ev.observe() if form in (GFT_FORM_MOMENTUM_F64, GFT_FORM_MOMENTUM_F32) else None
    rs, rv = G.grf_state(n, d, 0x1234_5678_9ABC, rms=0.1, slope=-1.5)
    err = np.abs(gs - rs).max()
    print(f"form {form}: max |GPU - CPU| = {err:.2e}")
    assert it == 0 and err <= tol
    assert np.array_equal(gv, rv)
    if obs is not None:
        assert obs.energy[2] == 0.0 or abs(obs.energy[3]) < 1e-30   # no scalar momentum at t = 0
        assert obs.energy[5] == 0.0                                   # E = 0


def test_grf_two_doublets_and_cutoff(rt):
    n, d = (16, 12, 18), (0.5, 0.5, 0.5)
    desc = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_MOMENTUM_F64, n=n, d=d, pot2=TWO_HDM_POTENTIAL)
    with rt.Evolver(desc) as ev:
        ev.init_grf(seed=99, rms=0.1, slope=-1.0, kmax=2.0, vevs=(0.8, 0.6))
        gs, _, _ = ev.export_two_slot(np.float64)
    rs, _ = G.grf_state(n, d, 99, rms=0.1, slope=-1.0, kmax=2.0, vevs=(0.8, 0.6), ndoublets=2)
    err = np.abs(gs - rs).max()
    print(f"2HDM: max |GPU - CPU| = {err:.2e}")
    assert err <= 1e-12


def test_grf_rejects_unsupported_sides(rt):
    with rt.Evolver(make_desc(form=GFT_FORM_MOMENTUM_F64, n=(10, 8, 8))) as ev:
        with pytest.raises(rt.GftError, match="2\\^a 3\\^b"):
            ev.init_grf(seed=1)


def test_grf_256_cube_statistics(rt):
    """The 256^3 workload's state from the GPU generator: rms, mean and the
    vev as requested (checked through the FP64 observables)."""
    n = (256, 256, 256)
    with rt.Evolver(make_desc(form=GFT_FORM_MOMENTUM_F32, n=n)) as ev:
        ev.init_grf(seed=20261203, rms=0.1, slope=-1.0)
        w = ev.wilson_loops()
    # <|phi|^2> = v^2 + rms^2 (zero-mean fluctuations), to the rounding of the sums.
    assert abs(w.phi_sq - 1.01) < 1e-9 and w.w11_su2 == 1.0


# gft_init_gaussian: the same white noise as the GRF, unfiltered, sigma per
# component around the vev (the 32^3 workload's initial state).
@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-14), (GFT_FORM_TWO_SLOT_F64, 1e-14),
                                      (GFT_FORM_MOMENTUM_F32, 1e-8), (GFT_FORM_TWO_SLOT_F32, 7e-8)],
                         ids=["mom64", "two64", "mom32", "two32"])
def test_gaussian_matches_restatement(rt, form, tol):
    n = (32, 32, 32)
    desc = make_desc(form=form, n=n)
    with rt.Evolver(desc) as ev:
        ev.init_gaussian(seed=20261203, sigma=0.05, vevs=(1.0, 1.0))
        gs, gv, it = ev.export_two_slot(np.float64)
        if form in (GFT_FORM_MOMENTUM_F64, GFT_FORM_MOMENTUM_F32):
            o = ev.observe()
            assert o.energy[3] == 0.0 and o.energy[5] == 0.0      # P = 0, E = 0
    rs, rv = G.gaussian_state(n, 20261203, sigma=0.05)
    err = np.abs(gs - rs).max()
    print(f"form {form}: max |GPU - CPU| = {err:.2e}")
    assert it == 0 and err <= tol
    assert np.array_equal(gv, rv) and np.array_equal(gs[0], gs[1])
    # iid N(vev, sigma^2) per component: sample mean and standard deviation of
    # 32768 values within 4 standard errors, no correlation between components.
    x = gs[1] - np.array([0.0, 0.0, 1.0, 0.0])
    se = 0.05 / np.sqrt(x.shape[0])
    assert np.all(np.abs(x.mean(axis=0)) < 4 * se)
    assert np.all(np.abs(x.std(axis=0) - 0.05) < 4 * 0.05 / np.sqrt(2 * x.shape[0]))
    c = np.corrcoef(x.T)
    assert np.abs(c - np.eye(4)).max() < 4 / np.sqrt(x.shape[0])


def test_gaussian_two_doublets(rt):
    n = (16, 12, 18)
    desc = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_TWO_SLOT_F64, n=n, pot2=TWO_HDM_POTENTIAL)
    with rt.Evolver(desc) as ev:
        ev.init_gaussian(seed=7, sigma=0.1, vevs=(0.8, 0.6))
        gs, _, _ = ev.export_two_slot(np.float64)
    rs, _ = G.gaussian_state(n, 7, sigma=0.1, vevs=(0.8, 0.6), ndoublets=2)
    assert np.abs(gs - rs).max() <= 1e-14
