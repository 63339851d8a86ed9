"""Prepares run directories for the reference binaries and writes their
positional .cfg files.

TEST INFRASTRUCTURE ONLY (drives oracle/_ref/gft_reference and
oracle/_ref/gft_ref_harness).

The reference reads each .cfg positionally: it skips a header, then reads the
value(s) after each ':' in order (e.g. Core/Source/Lattice.cpp:47-143,
Core/Source/Model.cpp:173-221).  oracle/Makefile compiles the reference with
config and data paths relative to the working directory and keeps the shipped
files in oracle/_ref/cfg as a template; `prepare(root)` copies the template to
root/cfg, and set_values() rewrites the values of the k-th value line after
the '=====' separator, leaving the description text untouched.
"""
from __future__ import annotations

import os
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATE = os.path.join(HERE, "_ref", "cfg")


def prepare(root: str) -> str:
    """Fresh run directory: root/cfg (copy of the template) and root/data/Continual."""
    dst = os.path.join(root, "cfg")
    if os.path.exists(dst):
        shutil.rmtree(dst)
    shutil.copytree(TEMPLATE, dst)
    for dp, _, files in os.walk(dst):
        for f in files:
            os.chmod(os.path.join(dp, f), 0o644)
    os.makedirs(os.path.join(root, "data", "Continual"), exist_ok=True)
    return root


def _fmt(v):
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, float):
        return repr(v)
    return str(v)


def set_values(root: str, lib: str, name: str, values: list) -> str:
    """values[k] (a scalar or a sequence) replaces the k-th value line."""
    path = os.path.join(root, "cfg", lib, "Config", name)
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
        txt = " ".join(_fmt(x) for x in v) if isinstance(v, (list, tuple)) else _fmt(v)
        head = lines[i].split(":", 1)[0]
        lines[i] = f"{head}:    {txt}"
        k += 1
    if k != len(values):
        raise ValueError(f"{path}: {len(values)} values for {k} value lines")
    with open(path, "w") as f:
        f.write("\n".join(lines))
    return path


def write_model(root: str, model: int, g: float, gp: float, *, lam=0.13, eta_sqr=1.0, pot2=None,
                damping=0.0, nt_damped=0, tol=1e-9, gen=False, gpp=0.5) -> str:
    """Model.cfg plus the potential / gradient / Wilson-loop configs; returns the
    Model.cfg path (relative to root, as the harness opens it from there).

    model 0 = SM (SO_N, SM gradient, SM loops); 1 = 2HDM with gaugeNum 4
    (Double_SO_N, TwoHDM gradient, TwoHDM loops, quaternion links); 2 = global
    O(N) model (SO_N, or Double_SO_N when pot2 is given; global gradient; null
    Wilson loop).
    pot2: dict of Double_SO_N parameters (m1 m2 l1 l2 l3 m12r l4p5 l6r l7r m12i
    l4m5 l5i l6i l7i break_to_son break_to_uhalfn pot_norm pot_offset).
    gen: the 2HDM's generator-representation flag (TwoHDM.cfg); the SM takes its
    representation from the vector component count (write_lattice's nvd).
    """
    if model == 2:
        ids = (2 if pot2 else 1, 1, 0)
        if not pot2:
            set_values(root, "Potentials", "SO_N.cfg", [lam, eta_sqr])
    elif model == 0:
        ids = (1, 2, 1)
        set_values(root, "Potentials", "SO_N.cfg", [lam, eta_sqr])
        set_values(root, "WilsonLoops", "StandardModel.cfg", [g, gp, tol])
        set_values(root, "Gradients", "StandardModel.cfg", [tol])
    else:
        ids = (2, 3, 2)
    if pot2:
        q = pot2
        set_values(root, "Potentials", "Double_SO_N.cfg", [
            4, int(q.get("pot_norm", 0)), q.get("pot_offset", 0.0),
            (q["m1"], q["m2"]), (q["l1"], q["l2"]), q["l3"],
            int(q.get("break_to_son", 1)), q["m12r"], q["l4p5"], (q["l6r"], q["l7r"]),
            int(q.get("break_to_uhalfn", 1)), q["m12i"], (q["l4m5"], q["l5i"]), (q["l6i"], q["l7i"]),
        ])
    if model == 1:
        set_values(root, "WilsonLoops", "TwoHDM.cfg", [g, gp, gpp, int(gen), tol])
        set_values(root, "Gradients", "TwoHDM.cfg", [tol])
    set_values(root, "Core", "Model.cfg", [ids[0], ids[1], ids[2], float(damping), nt_damped, 0, 0.0])
    return os.path.join("cfg", "Core", "Config", "Model.cfg")


def write_lattice(root: str, n, d, nt, dt, ns, *, ic_seed=42, ic_range=(-1.0, 1.0), efreq=1, gfreq=1,
                  energy=True, gauss=True, progress=False, faces=(3, 3, 3, 3, 3, 3), nvd=5, output_fields=False):
    """Lattice.cfg, RandomUniform.cfg and analyser configs for a full program run
    with a RandomUniform scalar initial state and identity links.  faces: the
    six face boundary types (-x, +x, -y, +y, -z, +z): 1 fixed, 3 periodic.
    nvd: vector components per direction, 5 (quaternion) or 4 (generator form).
    output_fields: switch the OutputFields analyser on (its own options in
    OutputFields.cfg)."""
    set_values(root, "Core", "Lattice.cfg", [
        tuple(n), tuple(d), nt, dt, ns, nvd, 0, 3.0, (1, 0), int(output_fields), int(energy), int(gauss), tuple(faces),
        int(progress), max(nt, 1),
    ])
    set_values(root, "InitialConditions", "RandomUniform.cfg", [ic_seed, tuple(ic_range), 1])
    set_values(root, "Analysers", "Energy.cfg", [
        "global_quantities.dat", 1, 1, 1, 1, 1, 1, efreq,
        "Continual/local_quantities", 0, 0, 0, 0, 0, 0, efreq,
    ])
    set_values(root, "Analysers", "GaugeCondition.cfg", [
        "global_quantities_gauge.dat", 1, 1, gfreq, "Continual/local_quantities_gauge", 0, gfreq,
    ])
    set_values(root, "Analysers", "OutputFields.cfg", [0, 0, 0, 0, 10, "initial_fields.dat",
                                                       "Continual/fields", "final_fields.dat"])
