"""GPU parity of the 2HDM with a Higgs-family gauge field (gaugeNum 5: an extra
U(1) under which the two doublets carry opposite charge; gaugeNum 7: an extra
SU(2) acting on the doublet index; Gradients/Source/TwoHDM.cpp,
WilsonLoops/Source/TwoHDM.cpp) in the two-slot forms, which follow the
reference's algorithm (the momentum forms: test_gpu_family_momentum.py):

  * two-slot FP32 against the fields the reference's own code wrote from the
    same state (harness_2hdm_g<5|7>[_gen].npz): a few float ulps;
  * two-slot FP64 against the FP64 oracle: <= 1e-12 relative (fields),
    energies <= 1e-10 relative, all gaugeNum Gauss integrals to 1e-11 of the
    cancelling terms;
  * a single doublet refuses a Higgs-family field with GFT_EINVAL.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64, GFT_MODEL_2HDM,
                                 GFT_MODEL_SM, TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L
from tests import oracle_lib as O
from tests.test_gpu_generator_form import quat_to_gen

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [(5, False), (5, True), (7, False), (7, True)]
IDS = ["g5", "g5-gen", "g7", "g7-gen"]


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def family_state(n, gauge_num, gen, seed=17):
    """2HDM near-vacuum state; links [y, SU(2), family part] per direction, the
    family part a small phase (5) or a near-identity SU(2) (7), with small
    t - dt offsets; converted to the generator form when gen is set."""
    s, v = L.near_vacuum_state(n, 8, seed=seed, amp=0.1, mom=0.01, link_eps=0.08, vev=0.7)
    N = s.shape[1]
    rng = np.random.default_rng(seed + 1)
    v = v.reshape(2, N, 3, 5)
    if gauge_num == 5:
        y2 = rng.normal(size=(N, 3, 1)) * 0.08
        fam = np.stack([y2, y2 + rng.normal(size=(N, 3, 1)) * 0.01])[::-1]
    else:
        _, v2 = L.near_vacuum_state(n, 4, seed=seed + 2, amp=0.1, mom=0.01, link_eps=0.08)
        fam = v2.reshape(2, N, 3, 5)[..., 1:]
    if gen:
        sm = quat_to_gen(v.reshape(2, N, 15)).reshape(2, N, 3, 4)
        if gauge_num == 7:
            q = np.concatenate([np.zeros(fam.shape[:-1] + (1,)), fam], axis=-1)   # [0, c0'..c3'] per direction
            fam = quat_to_gen(q.reshape(2, N, 15)).reshape(2, N, 3, 4)[..., 1:]
        v = np.concatenate([sm, fam], axis=-1)
    else:
        v = np.concatenate([v, fam], axis=-1)
    return s, v.reshape(2, N, -1).copy()


@pytest.mark.parametrize("gauge_num,gen", CASES, ids=IDS)
def test_twoslot_f32_family_gauge_matches_reference_golden(rt, gauge_num, gen):
    tag = f"2hdm_g{gauge_num}" + ("_gen" if gen else "")
    g = np.load(os.path.join(GOLD, f"harness_{tag}.npz"))
    d = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_TWO_SLOT_F32, n=(4, 5, 6), pot2=TWO_HDM_POTENTIAL,
                  gen_rep=int(gen), gauge_num=gauge_num, gpp=0.5)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(g["scalar0"], g["vec0"], 0)
        for k in range(6):
            rows = ev.evolve(1, energy_freq=3, gauss_freq=3)
            if rows and rows[0].has_gauss:
                assert rows[0].num_eq == gauge_num
            if f"scalar_{k}" in g:
                s, v, _ = ev.export_two_slot(np.float32)
                assert v.shape == g[f"vec_{k}"].shape
                ds = np.abs(s - g[f"scalar_{k}"]).max()
                dv = np.abs(v - g[f"vec_{k}"]).max()
                print(f"{tag} step {k}: max |GPU - reference| scalar {ds:.2e} links {dv:.2e}")
                assert ds <= 2e-6 and dv <= 4e-6


@pytest.mark.parametrize("gauge_num,gen", CASES, ids=IDS)
def test_twoslot_f64_family_gauge_matches_oracle(rt, gauge_num, gen):
    n = (10, 8, 12)
    d = make_desc(model=GFT_MODEL_2HDM, form=GFT_FORM_TWO_SLOT_F64, n=n, pot2=TWO_HDM_POTENTIAL, gen_rep=int(gen),
                  gauge_num=gauge_num, gpp=0.5)
    s, v = family_state(n, gauge_num, gen)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(30, energy_freq=10, gauss_freq=10)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 30 and gv.shape == v.shape
    _, orows, _ = O.evolve(d, s, v, 30, energy_freq=10, gauss_freq=10, nthreads=8)
    rs, rv = _rel(gs, s), _rel(gv, v)
    print(f"gaugeNum {gauge_num} gen {gen}: FP64 two-slot vs oracle: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= 1e-12 and rv <= 1e-12
    assert len(rows) == len(orows) == 3
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter and a.num_eq == b.num_eq == gauge_num
        np.testing.assert_allclose(np.array(a.energy[:]), np.array(b.energy[:]), rtol=1e-10,
                                   atol=1e-11 * abs(b.energy[0]))
        gi, go = np.array(a.gauss_int[:gauge_num]), np.array(b.gauss_int[:gauge_num])
        sc = np.array(b.gauss_scale[:gauge_num])
        print(f"  step {a.time_iter}: |G| {gi}, oracle {go}")
        assert (np.abs(gi - go) <= 1e-11 * sc).all()
        np.testing.assert_allclose(np.array(a.gauss_scale[:gauge_num]), sc, rtol=1e-10)


@pytest.mark.parametrize("form", [GFT_FORM_TWO_SLOT_F64, GFT_FORM_MOMENTUM_F32], ids=["two_slot", "momentum"])
def test_family_gauge_needs_two_doublets(rt, form):
    for g in (5, 7):
        d = make_desc(model=GFT_MODEL_SM, form=form, n=(8, 8, 8), gauge_num=g)
        with pytest.raises(rt.GftError) as e:
            rt.Evolver(d)
        assert e.value.code == -1 and "two doublets" in str(e.value)
