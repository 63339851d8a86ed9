"""The 32^3 SM workload as specified, end to end on the GPU: phi at the vev plus
iid Gaussian fluctuations of sigma = 0.05 per real component, seeded and
generated on the device (gft_init_gaussian), identity links, E = 0, zero
scalar momenta (so the Gauss law holds exactly at t = 0); 200 leapfrog steps
with the energy, the Gauss-law violation and the average plaquette logged at
every step.

  * FP64 two-slot and FP64 momentum forms against the CPU oracle run from the
    same (exported) initial state: fields, and the energy and Gauss rows of
    every step (Energy.cpp:117-220, GaugeCondition.cpp:117-226);
    the plaquettes are measured in the momentum forms (gft_wilson_loops);
  * the FP32-bulk momentum form (the fast path) against the FP64 momentum
    form: energy columns, the Gauss-law violation and the 1x1 plaquettes at
    every step."""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_FORM_TWO_SLOT_F64, make_desc
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

N3 = (32, 32, 32)
STEPS = 200
SEED = 20261204


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _gauss_rel(r):
    return max(r.gauss_int[m] / r.gauss_scale[m] for m in range(4) if r.gauss_scale[m] > 0)


def _run(rt, form, state=None):
    """200 steps, one row per step (energy before, Gauss after its update) and
    the plaquettes after every step.  Returns (initial two-slot state, rows,
    plaquettes [step][w11_su2, w11_u1], final fields)."""
    with rt.Evolver(make_desc(form=form, n=N3)) as ev:
        if state is None:
            ev.init_gaussian(SEED, sigma=0.05)
        else:
            ev.import_two_slot(state[0], state[1], 0)
        s0, v0, it0 = ev.export_two_slot(np.float64)
        assert it0 == 0
        rows, plaq = [], []
        for _ in range(STEPS):
            rows += ev.evolve(1, energy_freq=1, gauss_freq=1)
            if form != GFT_FORM_TWO_SLOT_F64:   # the loops are measured in the momentum forms
                w = ev.wilson_loops()
                plaq.append((w.w11_su2, w.w11_u1))
        s1, v1, it1 = ev.export_two_slot(np.float64)
    assert it1 == STEPS and len(rows) == STEPS
    return (s0, v0), rows, np.array(plaq), (s1, v1)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_config0_fp64_forms_match_oracle_every_step(rt):
    init, rows2, _, fin2 = _run(rt, GFT_FORM_TWO_SLOT_F64)
    s, v = init[0].copy(), init[1].copy()
    # The initial state: vev plus sigma = 0.05 per component, slots equal, at rest.
    assert np.array_equal(s[0], s[1]) and np.array_equal(v[0], v[1])
    dev = s[0] - np.array([0.0, 0.0, 1.0, 0.0])
    assert abs(dev.std() - 0.05) < 0.002 and abs(dev.mean()) < 0.002
    _, orows, _ = O.evolve(make_desc(form=GFT_FORM_TWO_SLOT_F64, n=N3), s, v, STEPS, energy_freq=1, gauss_freq=1,
                           nthreads=0)
    assert len(orows) == STEPS
    rs, rv = _rel(fin2[0], s), _rel(fin2[1], v)
    print(f"two-slot FP64 vs oracle after {STEPS} steps: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= 1e-12 and rv <= 1e-12
    init_m, rowsm, _, finm = _run(rt, GFT_FORM_MOMENTUM_F64, init)
    rms, rmv = _rel(finm[0], s), _rel(finm[1], v)
    print(f"momentum FP64 vs oracle after {STEPS} steps: scalar {rms:.2e}, links {rmv:.2e}")
    assert rms <= 1e-11 and rmv <= 1e-11
    worst = 0.0
    for rows, tol_e in ((rows2, 1e-10), (rowsm, 1e-9)):
        for a, b in zip(rows, orows):
            assert a.time_iter == b.time_iter and a.has_energy and a.has_gauss
            ea, eb = np.array(a.energy[:6]), np.array(b.energy[:6])
            worst = max(worst, np.abs(ea - eb).max() / abs(eb[0]))
            # The magnetic and electric densities 1 - cos(theta), 1 - P0 cancel for small
            # angles, so those columns agree to their evaluation order: 5e-11 of the total.
            np.testing.assert_allclose(ea, eb, rtol=tol_e, atol=5e-11 * abs(eb[0]))
            # |G| is at rounding level in FP64; compare on the scale of its terms.
            assert np.abs(np.array(a.gauss_int[:4]) - np.array(b.gauss_int[:4])).max() <= 1e-11 * max(b.gauss_scale[:4])
    print(f"energy columns vs oracle, every step: max {worst:.2e} of the total")


def test_config0_fp32_fast_path_every_step(rt):
    init, rows64, plaq64, _ = _run(rt, GFT_FORM_MOMENTUM_F64)
    _, rows32, plaq32, _ = _run(rt, GFT_FORM_MOMENTUM_F32, init)
    de = max(np.abs(np.array(a.energy[:6]) - np.array(b.energy[:6])).max() / abs(b.energy[0])
             for a, b in zip(rows32, rows64))
    g32 = max(_gauss_rel(r) for r in rows32)
    dp = np.abs(plaq32 - plaq64).max()
    e0, e1 = rows64[0].energy[0], rows64[-1].energy[0]
    print(f"FP32 vs FP64 over {STEPS} steps: energy columns {de:.2e} of the total, plaquettes {dp:.2e}, "
          f"FP32 Gauss violation max {g32:.2e}; FP64 energy drift {(e1 - e0) / abs(e0):.2e}")
    assert de <= 1e-6
    assert g32 <= 1e-6
    assert dp <= 1e-7
