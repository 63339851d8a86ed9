"""Long-run validation workloads (FP32-bulk against the FP64 evolver on the GPU).

    python -m gftevolver_amd.validate sm256   [--steps 10000] [--every 100]
    python -m gftevolver_amd.validate twohdm192 [--steps 5000] [--every 50]

sm256: SM 256^3, GRF Higgs state (P(k) ~ 1/k, k <= pi/(2dx), rms 0.1 v,
links identity, E = 0), dx = 0.5, dt = 0.05.  The FP32 and FP64 momentum
forms run the same steps from the same state; reported: energy drift of each,
the largest Gauss-law violation over the run (integrated |G| over the
magnitude of the cancelling terms, per generator), the field-level difference
FP32 - FP64 at the end, and the observables of both.

twohdm192: 2HDM 192^3, both doublets GRFs of rms 0.1 around vevs 0.8 and 0.6,
m11^2 = m22^2 = -0.5, m12^2 = 0.1, lambda1 = lambda2 = 0.26, lambda3 = 0.2,
lambda4 = -0.1, lambda5 = 0.05; 1x1 and 2x2 Wilson loops (SU(2) and U(1)) and
the magnetic energy every `every` steps, FP32 against FP64.

Each prints one progress line per measurement and a JSON summary at the end.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from ._abi import GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_MODEL_2HDM, TWO_HDM_POTENTIAL, make_desc


def log(*a):
    print(*a, flush=True)


def _gauss_rel(row):
    return max(row.gauss_int[m] / row.gauss_scale[m] for m in range(4) if row.gauss_scale[m] > 0)


def _loops(ev):
    w = ev.wilson_loops()
    return [w.w11_su2, w.w11_u1, w.w22_su2, w.w22_u1, w.phi_sq]


def _run(rt, desc, init, steps, every, tag):
    """Evolves in chunks of `every` steps; after each chunk measures the energy,
    the Gauss law (gft_observe) and the Wilson loops of the current state."""
    series = []
    with rt.Evolver(desc) as ev:
        init(ev)
        e0 = ev.observe().energy[:]
        series.append({"step": 0, "energy": list(e0), "loops": _loops(ev), "gauss_rel": 0.0})
        done, wall = 0, 0.0
        while done < steps:
            k = min(every, steps - done)
            t0 = time.perf_counter()
            ev.evolve(k)
            ev.sync()
            wall += time.perf_counter() - t0
            done += k
            o = ev.observe()
            series.append({"step": done, "energy": list(o.energy[:]), "loops": _loops(ev),
                           "gauss_rel": _gauss_rel(o)})
            if len(series) % 10 == 1 or done == steps:
                log(f"{tag}: step {done}, E {o.energy[0]:.9e} (drift {(o.energy[0] - e0[0]) / abs(e0[0]):+.2e}), "
                    f"Gauss rel {series[-1]['gauss_rel']:.2e}, loops {np.round(series[-1]['loops'], 7).tolist()}")
        out = ev.export_two_slot(np.float32)
    return series, out, wall


def _field_diff(a, b, ns):
    """Max |FP32 - FP64| of the current slot: scalar, U(1) phase (mod 2 pi), quaternion."""
    (s32, v32, it), (s64, v64, _) = a, b
    cur = (it + 1) % 2
    ds = float(np.abs(s32[cur] - s64[cur]).max())
    vv = (v32[cur] - v64[cur]).reshape(-1, 3, 5)
    dy = float(np.abs((vv[..., 0] + np.pi) % (2 * np.pi) - np.pi).max())
    dq = float(np.abs(vv[..., 1:]).max())
    return {"scalar_abs": ds, "scalar_rel": ds / float(np.abs(s64[cur]).max()), "u1_phase_abs": dy,
            "su2_abs": dq}


def _summary(name, steps, s32, s64, o32, o64, w32, w64, nsites):
    e32, e64 = s32[-1]["energy"][0], s64[-1]["energy"][0]
    out = {
        "workload": name, "steps": steps, "sites": nsites,
        "fp32_site_updates_per_s": nsites * steps / w32, "fp64_site_updates_per_s": nsites * steps / w64,
        "energy_drift_fp32": (e32 - s32[0]["energy"][0]) / abs(s32[0]["energy"][0]),
        "energy_drift_fp64": (e64 - s64[0]["energy"][0]) / abs(s64[0]["energy"][0]),
        "energy_fp32_vs_fp64_rel": abs(e32 - e64) / abs(e64),
        "gauss_rel_max_fp32": max(r["gauss_rel"] for r in s32),
        "gauss_rel_max_fp64": max(r["gauss_rel"] for r in s64),
        "gauss_rel_final_fp32": s32[-1]["gauss_rel"],
        "field_diff_fp32_vs_fp64": _field_diff(o32, o64, 4),
        "loops_names": ["w11_su2", "w11_u1", "w22_su2", "w22_u1", "phi_sq"],
        "loops_final_fp32": s32[-1]["loops"], "loops_final_fp64": s64[-1]["loops"],
        "loops_max_abs_diff_over_run": np.abs(np.array([r["loops"] for r in s32])
                                              - np.array([r["loops"] for r in s64])).max(axis=0).tolist(),
        "magnetic_energy_max_rel_diff_over_run": max(abs(a["energy"][4] - b["energy"][4]) / max(abs(b["energy"][0]), 1e-300)
                                                     for a, b in zip(s32, s64)),
        "electric_energy_max_rel_diff_over_run": max(abs(a["energy"][5] - b["energy"][5]) / max(abs(b["energy"][0]), 1e-300)
                                                     for a, b in zip(s32, s64)),
    }
    return out


def sm256(rt, steps=10000, every=100, n=256):
    N = n ** 3
    init = lambda ev: ev.init_grf(seed=20261203, rms=0.1, slope=-1.0)   # GRF state generated on the GPU
    s32, o32, w32 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F32, n=(n, n, n)), init, steps, every, "fp32")
    s64, o64, w64 = _run(rt, make_desc(form=GFT_FORM_MOMENTUM_F64, n=(n, n, n)), init, steps, every, "fp64")
    return _summary(f"SM {n}^3 GRF", steps, s32, s64, o32, o64, w32, w64, N)


def twohdm192(rt, steps=5000, every=50, n=192):
    N = n ** 3
    init = lambda ev: ev.init_grf(seed=20261204, rms=0.1, slope=-1.0, vevs=(0.8, 0.6))
    mk = lambda form: make_desc(model=GFT_MODEL_2HDM, form=form, n=(n, n, n), pot2=TWO_HDM_POTENTIAL)
    s32, o32, w32 = _run(rt, mk(GFT_FORM_MOMENTUM_F32), init, steps, every, "fp32")
    s64, o64, w64 = _run(rt, mk(GFT_FORM_MOMENTUM_F64), init, steps, every, "fp64")
    out = _summary(f"2HDM {n}^3 GRF", steps, s32, s64, o32, o64, w32, w64, N)
    out["series_every"] = every
    out["loops_series_fp32"] = [r["loops"] for r in s32]
    out["magnetic_series_fp32"] = [r["energy"][4] for r in s32]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=("sm256", "twohdm192"))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--every", type=int, default=None)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--out", type=str, default="")
    a = ap.parse_args(argv)
    from . import runtime as rt
    kw = {k: getattr(a, k) for k in ("steps", "every", "n") if getattr(a, k) is not None}
    res = (sm256 if a.workload == "sm256" else twohdm192)(rt, **kw)
    txt = json.dumps(res)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
