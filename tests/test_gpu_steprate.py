"""The step-rate driver (python -m gftevolver_amd.steprate) runs every model
and form it offers and reports a positive rate."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def test_steprate_models(capsys):
    from gftevolver_amd import steprate
    steprate.main(["--model", "sm", "--n", "16", "--steps", "3", "--form", "momentum_f32_plain"])
    steprate.main(["--model", "2hdm", "--gauge", "4", "5", "7", "--n", "16", "--steps", "3", "--launch", "16,16,3,4,2"])
    steprate.main(["--model", "2hdm", "--gauge", "5", "--n", "12", "--steps", "2", "--form", "momentum_f64"])
    lines = [l for l in capsys.readouterr().out.splitlines() if "ms per step" in l]
    assert len(lines) == 5
    for l in lines:
        ms = float(l.split(": ")[1].split(" ms")[0])
        assert ms > 0.0, l
