"""Classical-statistical ensembles of SM lattices on one GPU.

The ensemble workload: 64 independent SM lattices of uneven sizes
(16 x 32^3, 24 x 48^3, 16 x 64^3, 8 x 96^3; 14 450 688 sites in all), each a
Gaussian random field around the vev with P(k) ~ 1/k up to k_max = pi/(2 dx),
its own seed and an rms amplitude drawn uniformly from [0.05, 0.2]; couplings
and spacings of the 256^3 workload.  All members advance together in the
ensemble kernel (gft_ensemble_*), replayed from a two-step
hipGraph, with per-member tile shapes picked by the autotuner.

Observables per member, all FP64 reductions on the GPU:
  phi_sq    <|phi|^2> / v^2                       (gft_wilson_loops)
  e_frac    electric energy / total energy        (gft_observe)
  b_frac    magnetic energy / total energy
  plaq_su2  mean 1/2 tr of the 1x1 SU(2) Wilson loop
  plaq_u1   mean cos of the 1x1 U(1) Wilson loop
"""
from __future__ import annotations

import time

import numpy as np

from ._abi import GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, make_desc

OBSERVABLES = ("phi_sq", "e_frac", "b_frac", "plaq_su2", "plaq_u1")
GROUPS = ((16, 32), (24, 48), (16, 64), (8, 96))
SEED = 20261203


def members(groups=GROUPS, seed=SEED, rms_range=(0.05, 0.2)):
    """[(n, rms, seed)] of the ensemble workload; member i has seed `seed + i`."""
    rng = np.random.default_rng(seed)
    out = []
    for count, L in groups:
        for _ in range(count):
            out.append(((L, L, L), float(rng.uniform(*rms_range)), seed + len(out)))
    return out


def init_member(ev, rms, seed):
    """The member's GRF state, generated on the GPU (gft_init_grf)."""
    ev.init_grf(seed=seed, rms=rms, slope=-1.0)


def observe(ev):
    """The five ensemble observables of one member (an Evolver or an ensemble member view)."""
    o = ev.observe()
    w = ev.wilson_loops()
    tot = o.energy[0]
    return np.array([w.phi_sq, o.energy[5] / tot, o.energy[4] / tot, w.w11_su2, w.w11_u1])


def tune_tiles(ens, groups_idx, tiles=None, steps=20, passes=2):
    """Per-group tile shapes by coordinate descent on the whole ensemble's step
    time: for each group of equal-size members, try every tile shape with the
    other groups fixed and keep the fastest; `passes` sweeps over the groups.
    Every shape gives bitwise the same fields, so tuning advances the state but
    never changes results.  Returns {group: (bz, by)}."""
    from .tune import TILES
    tiles = tiles or TILES
    best = {g: tiles[0] for g in range(len(groups_idx))}
    for _ in range(passes):
        for g, idx in enumerate(groups_idx):
            timing = {}
            for t in tiles:
                for i in idx:
                    ens.set_launch(i, *t)
                ens.evolve(2)
                ens.evolve(steps)
                timing[t] = ens.last_step_ms()
            best[g] = min(timing, key=timing.get)
            for i in idx:
                ens.set_launch(i, *best[g])
    return best


def fixed_tiles(sizes, tiles):
    """{member shape: (bz, by)} for every shape in `sizes` from `tiles`, whose
    keys are member shapes (nx, ny, nz) or, for cubic members, the side."""
    given = {n: tiles.get(n, tiles.get(n[0]) if len(set(n)) == 1 else None) for n in sizes}
    missing = [n for n, t in given.items() if t is None]
    if missing:
        raise ValueError(f"no tile given for member sizes {missing}")
    return {n: tuple(t) for n, t in given.items()}


def run_fp32(rt, specs, steps, measure_every, tune=True, use_graph=True, log=None, tiles=None):
    """Advances the whole ensemble `steps` steps from t = 0.  Returns
    (observables [n_meas][n_members][5], timed seconds, avg kernel ms per step,
    tiles).  Timing excludes tuning and measurements: only the evolve calls.
    `tiles` ({size: (bz, by)}) fixes the tile of every member of each size and
    skips the tuner."""
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=n) for n, _, _ in specs]
    sizes = sorted({n for n, _, _ in specs})
    groups_idx = [[i for i, s in enumerate(specs) if s[0] == n] for n in sizes]
    given = fixed_tiles(sizes, tiles) if tiles is not None else None
    with rt.Ensemble(descs) as ens:
        if given is not None:
            for g, n in enumerate(sizes):
                for i in groups_idx[g]:
                    ens.set_launch(i, *given[n])
            tiles = {g: tuple(given[n]) for g, n in enumerate(sizes)}
        elif tune:
            # Tune on a throw-away copy of the state, then start afresh.
            for m, (_, rms, seed) in zip(ens.members, specs):
                init_member(m, rms, seed)
            tiles = tune_tiles(ens, groups_idx)
            if log:
                log(f"ensemble tiles per size {dict(zip(sizes, tiles.values()))}")
        for m, (_, rms, seed) in zip(ens.members, specs):
            init_member(m, rms, seed)
        obs = [[observe(m) for m in ens.members]]
        timed = 0.0
        kms = []
        done = 0
        while done < steps:
            k = min(measure_every, steps - done)
            t0 = time.perf_counter()
            ens.evolve(k, use_graph=use_graph)
            timed += time.perf_counter() - t0
            kms.append(ens.last_step_ms())
            done += k
            obs.append([observe(m) for m in ens.members])
    return np.array(obs), timed, float(np.mean(kms)), tiles


def run_fp64_sequential(rt, specs, steps, measure_every):
    """The same members one after another in the FP64 momentum form."""
    out = []
    for n, rms, seed in specs:
        with rt.Evolver(make_desc(form=GFT_FORM_MOMENTUM_F64, n=n)) as ev:
            init_member(ev, rms, seed)
            rows = [observe(ev)]
            done = 0
            while done < steps:
                k = min(measure_every, steps - done)
                ev.evolve(k)
                done += k
                rows.append(observe(ev))
        out.append(rows)
    return np.array(out).transpose(1, 0, 2)
