"""Writes the reference's positional .cfg files for one run.

TEST INFRASTRUCTURE ONLY (used to drive oracle/_ref/gft_reference and
oracle/_ref/gft_ref_harness in the build container).

The reference reads each .cfg positionally: it skips a header, then reads the
value(s) after each ':' in order (e.g. Core/Source/Lattice.cpp:47-143,
Core/Source/Model.cpp:173-221).  `oracle/Makefile` copies the shipped files into
oracle/_ref/cfg/<Lib>/Config; this module rewrites the values of the k-th value
line after the '=====' separator, leaving the description text untouched.
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
CFG = os.path.join(HERE, "_ref", "cfg")


def _fmt(v):
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, float):
        return repr(v)
    return str(v)


def set_values(lib: str, name: str, values: list) -> str:
    """values[k] (a scalar or a sequence) replaces the k-th value line."""
    path = os.path.join(CFG, lib, "Config", name)
    with open(path) as f:
        lines = f.read().split("\n")
    sep = next(i for i, l in enumerate(lines) if l.startswith("====="))
    k = 0
    for i in range(sep + 1, len(lines)):
        if ":" not in lines[i]:
            continue
        if k >= len(values):
            break
        v = values[k]
        if isinstance(v, (list, tuple)):
            txt = " ".join(_fmt(x) for x in v)
        else:
            txt = _fmt(v)
        head = lines[i].split(":", 1)[0]
        lines[i] = f"{head}:    {txt}"
        k += 1
    if k != len(values):
        raise ValueError(f"{path}: {len(values)} values for {k} value lines")
    with open(path, "w") as f:
        f.write("\n".join(lines))
    return path


def write_model(model: int, g: float, gp: float, *, lam=0.13, eta_sqr=1.0, pot2=None,
                damping=0.0, nt_damped=0, tol=1e-9) -> str:
    """Model.cfg plus the potential / gradient / Wilson-loop configs.

    model 0 = SM (SO_N, SM gradient, SM loops); 1 = 2HDM with gaugeNum 4
    (Double_SO_N, TwoHDM gradient, TwoHDM loops, quaternion links).
    pot2: dict of Double_SO_N parameters (m1 m2 l1 l2 l3 m12r l4p5 l6r l7r m12i
    l4m5 l5i l6i l7i break_to_son break_to_uhalfn pot_norm pot_offset).
    """
    if model == 0:
        ids = (1, 2, 1)
        set_values("Potentials", "SO_N.cfg", [lam, eta_sqr])
        set_values("WilsonLoops", "StandardModel.cfg", [g, gp, tol])
        set_values("Gradients", "StandardModel.cfg", [tol])
    else:
        ids = (2, 3, 2)
        q = pot2
        set_values("Potentials", "Double_SO_N.cfg", [
            4, int(q.get("pot_norm", 0)), q.get("pot_offset", 0.0),
            (q["m1"], q["m2"]), (q["l1"], q["l2"]), q["l3"],
            int(q.get("break_to_son", 1)), q["m12r"], q["l4p5"], (q["l6r"], q["l7r"]),
            int(q.get("break_to_uhalfn", 1)), q["m12i"], (q["l4m5"], q["l5i"]), (q["l6i"], q["l7i"]),
        ])
        set_values("WilsonLoops", "TwoHDM.cfg", [g, gp, 0.5, 0, tol])
        set_values("Gradients", "TwoHDM.cfg", [tol])
    return set_values("Core", "Model.cfg", [ids[0], ids[1], ids[2], float(damping), nt_damped, 0, 0.0])


def write_lattice(n, d, nt, dt, ns, *, ic_seed=42, ic_range=(-1.0, 1.0), efreq=1, gfreq=1,
                  energy=True, gauss=True, progress=False):
    """Lattice.cfg, RandomUniform.cfg and analyser configs for a full program run
    with a RandomUniform scalar initial state and identity links."""
    set_values("Core", "Lattice.cfg", [
        tuple(n), tuple(d), nt, dt, ns, 5, 0, 3.0, (1, 0), 0, int(energy), int(gauss), (3, 3, 3, 3, 3, 3),
        int(progress), max(nt, 1),
    ])
    set_values("InitialConditions", "RandomUniform.cfg", [ic_seed, tuple(ic_range), 1])
    set_values("Analysers", "Energy.cfg", [
        "global_quantities.dat", 1, 1, 1, 1, 1, 1, efreq,
        "Continual/local_quantities", 0, 0, 0, 0, 0, 0, efreq,
    ])
    set_values("Analysers", "GaugeCondition.cfg", [
        "global_quantities_gauge.dat", 1, 1, gfreq, "Continual/local_quantities_gauge", 0, gfreq,
    ])
    set_values("Analysers", "OutputFields.cfg", [0, 0, 0, 0, 10, "initial_fields.dat",
                                                 "Continual/fields", "final_fields.dat"])
