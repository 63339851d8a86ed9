"""The full-size workloads on the GPU, checked through size-independent
properties (the FP64 oracle on the CPU would take hours at these sizes):

  * SM at 256^3 (16.8M sites), Gaussian-random-field state (rms 0.1, slope -1,
    k_max = pi / (2 dx)), the workload's full 10^4 steps: the FP32-bulk form
    against the FP64 momentum form from the same state, every 1000 steps --
    every energy column, the Wilson loops and <|phi|^2> within 1e-5 relative
    (the ensemble-observable bar is 1e-4), the FP32 Gauss violation <= 1e-6 of
    its cancelling terms at every measurement, 10^4 steps included;
  * 2HDM at 192^3 (two doublets around vevs 0.8 and 0.6), the workload's full
    5000 steps with the 1x1 and 2x2 Wilson loops (SU(2) and U(1)) and the
    magnetic energy every 50 steps: FP32-bulk against FP64 within 1e-5
    relative, the magnetic energy within 1e-5 of the total energy (it starts at
    zero from identity links, so it is compared on the scale of the total).
  * the 2HDM with a Higgs-family U(1) or SU(2) gauge field (gaugeNum 5, 7) at
    128^3 over 1000 steps under a family-symmetric potential: FP32 against
    FP64 every 200 steps, energies within 1e-5 of the total, loops within
    1e-5, the FP32 Gauss violation of every generator <= 1e-6.
The FP64 momentum form is itself held to the FP64 CPU oracle at 1e-11
(tests/test_gpu_momentum.py, tests/test_gpu_family_momentum.py), so these tie
the full-size FP32 runs to it.
"""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_MODEL_2HDM, TWO_HDM_POTENTIAL,
                                 make_desc)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _loops(lp):
    return np.array([lp.w11_su2, lp.w11_u1, lp.w22_su2, lp.w22_u1, lp.phi_sq])


def _run(rt, desc, steps, every, vevs=(1.0, 1.0)):
    out = []
    with rt.Evolver(desc) as ev:
        ev.init_grf(20261203, rms=0.1, slope=-1.0, vevs=vevs)
        e0 = np.array(ev.observe().energy[:])
        for _ in range(steps // every):
            ev.evolve(every)
            obs = ev.observe()   # energies and, after a step, the Gauss law of the last step
            out.append((np.array(obs.energy[:]), _loops(ev.wilson_loops()), obs))
    return e0, out


def test_sm_256_fp32_matches_fp64_observables(rt):
    n = (256, 256, 256)
    e0, r32 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F32, n=n), 10000, 1000)
    f0, r64 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F64, n=n), 10000, 1000)
    assert len(r32) == len(r64) == 10
    assert np.abs(e0 - f0).max() <= 1e-6 * abs(f0[0])
    for k, ((e32, l32, g32), (e64, l64, g64)) in enumerate(zip(r32, r64)):
        de = np.abs(e32 - e64).max() / abs(e64[0])
        dl = np.abs(l32 - l64).max() / np.abs(l64).max()
        assert g32.has_gauss
        gs = max(g32.gauss_int[m] / g32.gauss_scale[m] for m in range(4))
        print(f"step {1000 * (k + 1)}: energy {e32[0]:.9e} vs {e64[0]:.9e} (col err {de:.1e}), loops err {dl:.1e}, "
              f"FP32 Gauss {gs:.1e}")
        assert de <= 1e-5 and dl <= 1e-5
        assert gs <= 1e-6


def test_2hdm_192_fp32_matches_fp64_wilson_loops(rt):
    n = (192, 192, 192)
    kw = dict(model=GFT_MODEL_2HDM, n=n, pot2=TWO_HDM_POTENTIAL)
    _, r32 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F32, **kw), 5000, 50, vevs=(0.8, 0.6))
    _, r64 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F64, **kw), 5000, 50, vevs=(0.8, 0.6))
    assert len(r32) == len(r64) == 100
    worst = np.zeros(2)
    for k, ((e32, l32, _), (e64, l64, _)) in enumerate(zip(r32, r64)):
        dl = np.abs(l32 - l64) / np.abs(l64)
        dm = abs(e32[4] - e64[4]) / abs(e64[0])
        worst = np.maximum(worst, [dl.max(), dm])
        if k % 10 == 9:
            print(f"step {50 * (k + 1)}: loops {l32} rel err {dl.max():.1e}, magnetic energy rel err {dm:.1e}")
        assert dl.max() <= 1e-5 and dm <= 1e-5, (50 * (k + 1), dl, dm)
    print(f"worst over 100 measurements: loops {worst[0]:.1e}, magnetic energy {worst[1]:.1e}")


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
def test_2hdm_family_128_fp32_matches_fp64(rt, G):
    from tests.test_gpu_family_momentum import FAMILY_SYMMETRIC_POTENTIAL
    n = (128, 128, 128)
    kw = dict(model=GFT_MODEL_2HDM, n=n, pot2=FAMILY_SYMMETRIC_POTENTIAL, gauge_num=G, gpp=0.5)
    _, r32 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F32, **kw), 1000, 200, vevs=(0.8, 0.6))
    _, r64 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F64, **kw), 1000, 200, vevs=(0.8, 0.6))
    assert len(r32) == len(r64) == 5
    for k, ((e32, l32, g32), (e64, l64, g64)) in enumerate(zip(r32, r64)):
        de = np.abs(e32 - e64).max() / abs(e64[0])
        dl = np.abs(l32 - l64).max() / np.abs(l64).max()
        assert g32.has_gauss and g32.num_eq == G
        gs = max(g32.gauss_int[m] / g32.gauss_scale[m] for m in range(G))
        g64s = max(g64.gauss_int[m] / g64.gauss_scale[m] for m in range(G))
        print(f"gaugeNum {G} step {200 * (k + 1)}: energy {e32[0]:.9e} vs {e64[0]:.9e} (col err {de:.1e}), "
              f"loops err {dl:.1e}, Gauss FP32 {gs:.1e} FP64 {g64s:.1e}")
        assert de <= 1e-5 and dl <= 1e-5
        assert gs <= 1e-6 and g64s <= 1e-12
