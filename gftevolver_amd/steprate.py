"""Step rate of one model on one GPU: a GRF initial state generated on the GPU,
then timed steps of the chosen form (HIP events on the library's stream).

    python -m gftevolver_amd.steprate --model 2hdm --gauge 5 --n 192 --steps 8
    python -m gftevolver_amd.steprate --model sm --n 256 --form momentum_f32_plain --launch 64,4,3,4,2

Several --gauge values run one after another in the same process (one line
each), which is how the counter runs of the 2HDM kernels drive them under a
profiler.  Couplings are the workloads' (SM: g = 0.65, g' = 0.35,
lambda = 0.13; 2HDM: the two-Higgs-doublet workload potential, g'' = 0.5
for a Higgs-family field), dx = 0.5, dt = 0.05.
"""
from __future__ import annotations

import argparse

from ._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN, GFT_FORM_MOMENTUM_F64, GFT_MODEL_2HDM,
                   GFT_MODEL_SM, TWO_HDM_POTENTIAL, make_desc)

FORMS = {"momentum_f32": GFT_FORM_MOMENTUM_F32, "momentum_f32_plain": GFT_FORM_MOMENTUM_F32_PLAIN,
         "momentum_f64": GFT_FORM_MOMENTUM_F64}


def step_ms(rt, model: str, gauge: int, n: int, form: str, steps: int, warmup: int, launch=None) -> float:
    """Average ms per step over `steps` steps after `warmup` untimed ones."""
    if model == "sm":
        d = make_desc(model=GFT_MODEL_SM, form=FORMS[form], n=(n, n, n))
        vevs = (1.0, 1.0)
    else:
        d = make_desc(model=GFT_MODEL_2HDM, form=FORMS[form], n=(n, n, n), pot2=TWO_HDM_POTENTIAL, gauge_num=gauge,
                      gpp=0.5)
        vevs = (0.8, 0.6)
    with rt.Evolver(d) as ev:
        ev.init_grf(seed=3, rms=0.1, slope=-1.0, vevs=vevs)
        if launch:
            ev.set_launch(*launch)
        if warmup:
            ev.evolve(warmup)
        ev.evolve(steps)
        return ev.last_step_ms()[0]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--model", choices=("sm", "2hdm"), default="sm")
    ap.add_argument("--gauge", type=int, nargs="+", default=[4], help="gaugeNum of the 2HDM: 4, 5 or 7")
    ap.add_argument("--n", type=int, default=256, help="lattice side")
    ap.add_argument("--form", choices=tuple(FORMS), default="momentum_f32")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--launch", type=str, default="", help="bz,by,order,min_waves,stage")
    a = ap.parse_args(argv)
    from . import runtime as rt
    launch = tuple(int(v) for v in a.launch.split(",")) if a.launch else None
    sites = a.n ** 3
    for g in (a.gauge if a.model == "2hdm" else [4]):
        ms = step_ms(rt, a.model, g, a.n, a.form, a.steps, a.warmup, launch)
        print(f"{a.model} gaugeNum {g} {a.form} {a.n}^3: {ms:.4f} ms per step, "
              f"{sites / (ms * 1e-3):.3e} site-updates/s", flush=True)


if __name__ == "__main__":
    main()
