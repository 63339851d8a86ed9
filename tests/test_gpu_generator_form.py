"""GPU parity of the generator form of the SU(2) links (gft_desc.gen_rep = 1):
the two-slot arrays hold 3 x [y, w1, w2, w3] per site, U = exp(i w.sigma)
(WilsonLoops/Source/StandardModel.cpp:78-157, 195-210; TwoHDM.cpp:215-242).

  * Two-slot FP32 (the reference's storage and order) against the fields the
    reference's own code wrote from the same state (harness_<model>_gen.npz):
    a few float ulps, as for the quaternion form.  The state holds links with
    |w| up to 3, where the inverse conversion takes its pi - asin branch.
  * Two-slot FP64 against the FP64 oracle: <= 1e-12 relative (fields),
    energies <= 1e-10 relative.
  * Momentum forms, which convert at import and export only (the inverse as
    atan2(|c|, c0), well conditioned everywhere) and step
    quaternions inside: against the FP64 oracle stepping the same group
    elements in the quaternion form, <= 1e-11 (FP64) / 2e-5 (FP32) relative,
    the quaternion form's tolerances.  Against the oracle stepping the
    generator form itself (the reference's algorithm) the FP64 bound is 1e-9:
    that form loses precision by 1/|c0| where a link passes |w| = pi/2
    (asin(|c|) at |c| -> 1); on this state a link reaches c0 = -5e-5 and the
    two oracle forms part by 2.5e-11 after 60 steps while a 1e-15 perturbation
    of the quaternion form grows only to 2e-14.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN, GFT_FORM_MOMENTUM_F64,
                                 GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64, GFT_MODEL_2HDM, GFT_MODEL_SM,
                                 TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def gen_state(n, nd=1, seed=12, amp=0.15, mom=0.02, link_eps=0.1):
    """Near-vacuum state with links in the generator form: w = angle x axis of
    the quaternion links of lattice_np.near_vacuum_state."""
    s, v = L.near_vacuum_state(n, 4 * nd, seed=seed, amp=amp, mom=mom, link_eps=link_eps,
                               vev=0.7 if nd == 2 else 1.0)
    return s, quat_to_gen(v)


def quat_to_gen(v):
    q = v.reshape(v.shape[:-1] + (3, 5))
    c = q[..., 2:]
    cm = np.linalg.norm(c, axis=-1, keepdims=True)
    ph = np.arctan2(cm, q[..., 1:2])
    w = np.where(cm > 0, ph * c / np.where(cm > 0, cm, 1), 0.0)
    out = np.concatenate([q[..., :1], w], axis=-1)
    return out.reshape(v.shape[:-1] + (12,))


def gen_to_quat(v):
    g = v.reshape(v.shape[:-1] + (3, 4))
    w = g[..., 1:]
    a = np.linalg.norm(w, axis=-1, keepdims=True)
    c = np.where(a > 0, np.sin(a) * w / np.where(a > 0, a, 1), 0.0)
    out = np.concatenate([g[..., :1], np.cos(a), c], axis=-1)
    return out.reshape(v.shape[:-1] + (15,))


def _quat_err(vg, vq):
    """Max difference between generator-form links and quaternion-form links of
    the same group elements, compared as quaternions (phases modulo 2 pi)."""
    qa = gen_to_quat(vg).reshape(vg.shape[:-1] + (3, 5))
    qb = vq.reshape(vq.shape[:-1] + (3, 5))
    dy = (qa[..., 0] - qb[..., 0] + np.pi) % (2 * np.pi) - np.pi
    return max(np.abs(dy).max(), np.abs(qa[..., 1:] - qb[..., 1:]).max())


def _link_err(va, vb):
    """Max difference of two generator-form link arrays: phases modulo 2 pi,
    SU(2) parts compared as group elements (quaternions)."""
    qa, qb = gen_to_quat(va).reshape(va.shape[:-1] + (3, 5)), gen_to_quat(vb).reshape(vb.shape[:-1] + (3, 5))
    dy = (qa[..., 0] - qb[..., 0] + np.pi) % (2 * np.pi) - np.pi
    return max(np.abs(dy).max(), np.abs(qa[..., 1:] - qb[..., 1:]).max())


@pytest.mark.parametrize("model,tag,pot2", [(GFT_MODEL_SM, "sm", None), (GFT_MODEL_2HDM, "2hdm", TWO_HDM_POTENTIAL)],
                         ids=["sm", "2hdm"])
def test_twoslot_f32_generator_form_matches_reference_golden(rt, model, tag, pot2):
    g = np.load(os.path.join(GOLD, f"harness_{tag}_gen.npz"))
    d = make_desc(model=model, form=GFT_FORM_TWO_SLOT_F32, n=(4, 5, 6), pot2=pot2, gen_rep=1)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(g["scalar0"], g["vec0"], 0)
        for k in range(6):
            ev.evolve(1, energy_freq=3, gauss_freq=3)
            if f"scalar_{k}" in g:
                s, v, _ = ev.export_two_slot(np.float32)
                assert v.shape[-1] == 12
                ds = np.abs(s - g[f"scalar_{k}"]).max()
                dv = np.abs(v - g[f"vec_{k}"]).max()
                print(f"{tag} step {k}: max |GPU - reference| scalar {ds:.2e} links {dv:.2e}")
                # A few float ulps at |w| <= 3.
                assert ds <= 2e-6 and dv <= 4e-6


@pytest.mark.parametrize("model,pot2", [(GFT_MODEL_SM, None), (GFT_MODEL_2HDM, TWO_HDM_POTENTIAL)],
                         ids=["sm", "2hdm"])
def test_twoslot_f64_generator_form_matches_oracle(rt, model, pot2):
    n = (10, 8, 12)
    nd = 2 if model == GFT_MODEL_2HDM else 1
    d = make_desc(model=model, form=GFT_FORM_TWO_SLOT_F64, n=n, pot2=pot2, gen_rep=1)
    s, v = gen_state(n, nd)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(30, energy_freq=10, gauss_freq=10)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 30 and gv.shape == v.shape
    _, orows, _ = O.evolve(d, s, v, 30, energy_freq=10, gauss_freq=10, nthreads=8)
    rs, rv = _rel(gs, s), _rel(gv, v)
    print(f"model {model}: FP64 two-slot generator form vs oracle: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= 1e-12 and rv <= 1e-12
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter
        np.testing.assert_allclose(np.array(a.energy[:]), np.array(b.energy[:]), rtol=1e-10,
                                   atol=1e-11 * abs(b.energy[0]))
        assert np.abs(np.array(a.gauss_int[:]) - np.array(b.gauss_int[:])).max() <= 1e-11 * max(b.gauss_scale[:])


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-11), (GFT_FORM_MOMENTUM_F32, 2e-5),
                                      (GFT_FORM_MOMENTUM_F32_PLAIN, 2e-5)], ids=["f64", "f32", "f32plain"])
def test_momentum_generator_form_matches_oracle(rt, form, tol):
    n = (16, 12, 20)
    d = make_desc(form=form, n=n, gen_rep=1)
    s, v = gen_state(n)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(60, energy_freq=20)
        gs, gv, _ = ev.export_two_slot(np.float64)
    assert gv.shape == v.shape
    sq, vq = s.copy(), gen_to_quat(v)
    O.evolve(make_desc(form=form, n=n), sq, vq, 60, nthreads=8)
    _, orows, _ = O.evolve(d, s, v, 60, energy_freq=20, nthreads=8)
    cur = (60 + 1) % 2
    rq = max(_rel(gs[cur], sq[cur]), _quat_err(gv[cur], vq[cur]))
    rs, rv = _rel(gs[cur], s[cur]), _link_err(gv[cur], v[cur])
    print(f"form {form}: after 60 steps vs the quaternion-form oracle {rq:.2e}; vs the generator-form oracle: "
          f"scalar {rs:.2e}, links {rv:.2e}")
    assert rq <= tol
    gtol = 1e-9 if form == GFT_FORM_MOMENTUM_F64 else tol
    assert rs <= gtol and rv <= gtol
    etol = 1e-9 if form == GFT_FORM_MOMENTUM_F64 else 1e-5
    for a, b in zip(rows, orows):
        err = np.abs(np.array(a.energy[:]) - np.array(b.energy[:])).max() / abs(b.energy[0])
        assert err <= etol


def test_momentum_generator_and_quaternion_forms_agree(rt):
    """The momentum form keeps quaternions inside; the generator form only
    changes the conversions at the boundary, so the same state given in both
    forms evolves to the same group elements (FP64: to the conversions' rounding)."""
    n = (12, 10, 14)
    s, vq = L.near_vacuum_state(n, 4, seed=21, amp=0.15, mom=0.02, link_eps=0.3)
    vg = quat_to_gen(vq)
    vq = gen_to_quat(vg)          # the exact same group elements in both forms
    out = {}
    for gen, v in ((0, vq), (1, vg)):
        d = make_desc(form=GFT_FORM_MOMENTUM_F64, n=n, gen_rep=gen)
        with rt.Evolver(d) as ev:
            ev.import_two_slot(s, v, 0)
            ev.evolve(25)
            out[gen] = ev.export_two_slot(np.float64)
    ds = np.abs(out[0][0] - out[1][0]).max()
    dq = _quat_err(out[1][1], out[0][1])
    print(f"quaternion vs generator form after 25 steps: scalar {ds:.2e}, links {dq:.2e}")
    assert ds <= 1e-12 and dq <= 1e-12


def test_grf_initial_state_generator_form_identity_links(rt):
    """gft_init_grf writes identity links: w = 0 in the generator form."""
    n = (8, 6, 12)
    for form in (GFT_FORM_TWO_SLOT_F64, GFT_FORM_MOMENTUM_F32):
        d = make_desc(form=form, n=n, gen_rep=1)
        with rt.Evolver(d) as ev:
            ev.init_grf(3, rms=0.1)
            s, v, _ = ev.export_two_slot(np.float64)
        assert v.shape == (2, int(np.prod(n)), 12)
        assert np.array_equal(v, np.zeros_like(v)), form
        assert np.abs(s).max() > 0
