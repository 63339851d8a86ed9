"""GPU parity of the momentum forms (the fast path) for the 2HDM with a
Higgs-family gauge field (gaugeNum 5: a U(1) under which the two doublets carry
opposite charge; gaugeNum 7: an SU(2) acting on the doublet index;
Gradients/Source/TwoHDM.cpp:72-141, 299-319, 401-424,
WilsonLoops/Source/TwoHDM.cpp:352-386, 496-552, 600-702).  The momentum forms
store P and E instead of a second time slot and reorder the reference's
arithmetic, so they are compared with the FP64 oracle within tolerances:

  * FP64 momentum form vs the FP64 oracle after 40 steps: <= 1e-11 relative
    (fields; phases modulo 2 pi), energies within 1e-9 of the total at every
    measured step;
  * FP32 (compensated and plain storage) vs the FP64 oracle: <= 2e-5, energies
    within 1e-5 of the total;
  * generator-form links at the boundary: the same as the generator-form
    oracle to 1e-9 (FP64) and 2e-5 (FP32);
  * the Gauss law from a constraint-satisfying hot state under a potential
    with the family symmetry (the workload potential's m12^2 and lambda_5
    terms break it, and then the family generators' constraint drifts in the
    reference's own algorithm too, 1e-4 after 200 steps): every generator <=
    1e-12 (FP64) and <= 1e-6 (FP32) of its cancelling terms over 200 steps;
    under the breaking potential the family Gauss integrals equal the
    oracle's (FP64: 1e-9 of the scale);
  * every launch shape and staging mode gives bitwise the same fields;
  * checkpoint / restore resumes bitwise;
  * generated initial states carry identity family links with no family
    electric field, and keep the Gauss law.
"""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN, GFT_FORM_MOMENTUM_F64,
                                 GFT_MODEL_2HDM, TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L
from tests import oracle_lib as O
from tests.test_gpu_family_gauge import family_state

pytestmark = pytest.mark.gpu
FORMS = [(GFT_FORM_MOMENTUM_F64, 1e-11, 1e-9), (GFT_FORM_MOMENTUM_F32, 2e-5, 1e-5),
         (GFT_FORM_MOMENTUM_F32_PLAIN, 2e-5, 1e-5)]
# A Double_SO_N potential of |Phi1|^2 + |Phi2|^2 and |Phi1|^2 |Phi2|^2 - |Phi1^dag Phi2|^2
# only (f1 + f2 and f1 f2 - D^2 - A^2): invariant under the family U(1) and SU(2) as well
# as SU(2) x U(1), with the D^2 and A^2 terms on (break_to_son, break_to_uhalfn).
FAMILY_SYMMETRIC_POTENTIAL = dict(m1=0.25, m2=0.25, l1=0.0325, l2=0.0325, l3=0.115, break_to_son=1,
                                  break_to_uhalfn=1, m12r=0.0, l4p5=-0.05, l6r=0.0, l7r=0.0, m12i=0.0,
                                  l4m5=-0.05, l5i=0.0, l6i=0.0, l7i=0.0, pot_norm=0, pot_offset=0.0)
FORM_IDS = ["f64", "f32", "f32plain"]


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _desc(G, form, n, gen=False, pot2=TWO_HDM_POTENTIAL):
    return make_desc(model=GFT_MODEL_2HDM, form=form, n=n, pot2=pot2, gauge_num=G, gpp=0.5, gen_rep=int(gen))


def _link_diff(va, vb, G, gen=False):
    """Link differences per component, the U(1) phases (y, and y' for gaugeNum
    5) modulo 2 pi: the momentum form wraps them into (-pi, pi]."""
    vd = (4 if gen else 5) + (1 if G == 5 else 3 if gen else 4)
    d = (va - vb).reshape(va.shape[:-1] + (3, vd)).copy()
    for k in ([0, vd - 1] if G == 5 else [0]):
        d[..., k] = (d[..., k] + np.pi) % (2 * np.pi) - np.pi
    return d


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
@pytest.mark.parametrize("form,tol,etol", FORMS, ids=FORM_IDS)
def test_family_momentum_matches_oracle(rt, G, form, tol, etol):
    n = (10, 8, 12)
    d = _desc(G, form, n)
    s, v = family_state(n, G, False)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(40, energy_freq=10, gauss_freq=20)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 40 and gv.shape == v.shape
    _, orows, _ = O.evolve(d, s, v, 40, energy_freq=10, gauss_freq=20, nthreads=8)
    cur = (40 + 1) % 2
    rs, rv = _rel(gs[cur], s[cur]), np.abs(_link_diff(gv[cur], v[cur], G)).max()
    print(f"gaugeNum {G} form {form}: vs the FP64 oracle after 40 steps: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= tol and rv <= tol
    assert len(rows) == len(orows) == 4
    # The workload potential breaks the family symmetry: the family generators'
    # Gauss integrals are physical (non-zero) and must equal the oracle's.
    gtol = 1e-9 if form == GFT_FORM_MOMENTUM_F64 else 2e-6
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter and a.has_gauss == b.has_gauss
        err = np.abs(np.array(a.energy[:]) - np.array(b.energy[:])).max() / abs(b.energy[0])
        print(f"  step {a.time_iter}: energy {a.energy[0]:.10e} vs {b.energy[0]:.10e} (rel col err {err:.1e})")
        assert err <= etol
        if a.has_gauss:
            gi, go = np.array(a.gauss_int[:G]), np.array(b.gauss_int[:G])
            sc = np.array(b.gauss_scale[:G])
            print(f"  step {a.time_iter}: |G| / scale {gi / sc}, oracle {go / sc}")
            assert (np.abs(gi - go) <= gtol * sc).all()


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-9), (GFT_FORM_MOMENTUM_F32, 2e-5)],
                         ids=["f64", "f32"])
def test_family_momentum_generator_form_matches_oracle(rt, G, form, tol):
    """Generator-form links [y, w, y' or w'] enter and leave through the
    conversions at the boundary; the generator-form oracle steps the reference's
    own representation (asin inverse: 1e-11-level conditioning in FP64)."""
    n = (8, 6, 10)
    d = _desc(G, form, n, gen=True)
    s, v = family_state(n, G, True, seed=23)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(30, energy_freq=10)
        gs, gv, _ = ev.export_two_slot(np.float64)
    assert gv.shape == v.shape
    _, orows, _ = O.evolve(d, s, v, 30, energy_freq=10, nthreads=8)
    cur = (30 + 1) % 2
    rs, rv = _rel(gs[cur], s[cur]), np.abs(_link_diff(gv[cur], v[cur], G, gen=True)).max()
    print(f"gaugeNum {G} generator form {form}: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= tol and rv <= tol
    etol = 1e-9 if form == GFT_FORM_MOMENTUM_F64 else 1e-5
    for a, b in zip(rows, orows):
        assert np.abs(np.array(a.energy[:]) - np.array(b.energy[:])).max() / abs(b.energy[0]) <= etol


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-12), (GFT_FORM_MOMENTUM_F32, 1e-6),
                                      (GFT_FORM_MOMENTUM_F32_PLAIN, 1e-6)], ids=FORM_IDS)
def test_family_momentum_gauss_law_conserved(rt, G, form, tol):
    """From E = 0, P = 0 the discrete Gauss law holds exactly for all G
    generators when the potential has the family symmetry; a wrong sign, charge
    or rotation in the family force or in its Gauss terms breaks it at O(1).
    Hot state, ragged lattice, unequal spacings."""
    n = (12, 10, 14)
    d = make_desc(model=GFT_MODEL_2HDM, form=form, n=n, d=(0.5, 0.45, 0.55), pot2=FAMILY_SYMMETRIC_POTENTIAL,
                  gauge_num=G, gpp=0.5)
    s, v = family_state(n, G, False, seed=31)
    s[0], v[0] = s[1], v[1]
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(200, gauss_freq=50)
    assert all(r.num_eq == G for r in rows)
    rel = np.array([[r.gauss_int[m] / r.gauss_scale[m] for m in range(G)] for r in rows])
    print(f"gaugeNum {G} form {form}: relative Gauss violation per generator, every 50 steps:\n{rel}")
    assert rel.max() <= tol
    assert (np.array([[r.gauss_scale[m] for m in range(G)] for r in rows]) > 0).all()


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
def test_family_momentum_launch_shapes_bitwise_identical(rt, G):
    """Tiles, block orders and the staging modes (LDS plane tile; shared
    SU(2) x U(1) plaquettes and covariant differences) change where values
    come from, never the arithmetic: ragged tiles in every mode."""
    n = (16, 12, 20)
    d = _desc(G, GFT_FORM_MOMENTUM_F32, n)
    s, v = family_state(n, G, False, seed=5)
    cfgs = ((64, 4, 2, 4, 1), (64, 4, 2, 4, 0), (256, 1, 2, 4, 1), (16, 16, 3, 4, 1), (32, 8, 1, 4, 0),
            (128, 2, 4, 4, 1), (16, 16, 0, 4, 0), (64, 4, 3, 4, 2), (16, 16, 0, 4, 2), (32, 8, 4, 4, 2),
            (128, 2, 1, 4, 2))
    out = {}
    for c in cfgs:
        with rt.Evolver(d) as ev:
            ev.set_launch(*c)
            ev.import_two_slot(s, v, 0)
            ev.evolve(7)
            out[c] = ev.export_two_slot(np.float64)
    ref = out[cfgs[0]]
    for c in cfgs[1:]:
        nd_ = int((out[c][0] != ref[0]).sum() + (out[c][1] != ref[1]).sum())
        print(f"launch {c}: {nd_} values differ from {cfgs[0]}")
        assert nd_ == 0, c


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-11), (GFT_FORM_MOMENTUM_F32, 2e-5)],
                         ids=["f64", "f32"])
def test_family_momentum_fixed_faces_and_damping_match_oracle(rt, G, form, tol):
    """Fixed faces (both y faces and both x faces) and the damped first steps
    (Model::update, Core/Source/Model.cpp:241-254) with the family fields: the
    family links and electric fields take the frozen-shell and damping paths
    of the kick and drift.  Fields vs the FP64 oracle after 30 steps (5 damped)
    within the periodic tolerances; the shell bitwise unchanged."""
    n = (8, 6, 10)
    faces = 4 | 8 | 1 | 2
    d = make_desc(model=GFT_MODEL_2HDM, form=form, n=n, d=(0.5, 0.45, 0.55), pot2=TWO_HDM_POTENTIAL,
                  gauge_num=G, gpp=0.5, fixed_faces=faces, damping=1.0, nt_damped=5)
    s, v = family_state(n, G, False, seed=23)
    m = O.frozen_mask(n, faces)
    s[0][m], v[0][m] = s[1][m], v[1][m]
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        base_s, base_v, _ = ev.export_two_slot(np.float64)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        ev.evolve(30)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 30
    O.evolve(d, s, v, 30, nthreads=8)
    cur = (30 + 1) % 2
    rs, rv = _rel(gs[cur], s[cur]), np.abs(_link_diff(gv[cur], v[cur], G)).max()
    print(f"gaugeNum {G} form {form}: fixed faces + damping vs the oracle: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= tol and rv <= tol
    assert np.array_equal(gs[:, m], base_s[:, m]) and np.array_equal(gv[:, m], base_v[:, m])


@pytest.mark.parametrize("G,form", [(5, GFT_FORM_MOMENTUM_F32), (7, GFT_FORM_MOMENTUM_F32),
                                    (7, GFT_FORM_MOMENTUM_F64)], ids=["g5-f32", "g7-f32", "g7-f64"])
def test_family_momentum_resume_is_bitwise(rt, G, form):
    """N steps in one go equal K steps, checkpoint, restore into a fresh
    context, N - K steps: the family arrays and their residuals are part of the
    checkpointed state."""
    n = (12, 10, 14)
    d = _desc(G, form, n)
    s, v = family_state(n, G, False, seed=41)
    N, K = 9, 4
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows_a = ev.evolve(N, energy_freq=2, gauss_freq=3)
        fa = ev.export_two_slot(np.float64)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows_b = ev.evolve(K, energy_freq=2, gauss_freq=3)
        ck = ev.checkpoint()
    with rt.Evolver(d) as ev:
        ev.restore(ck)
        rows_b += ev.evolve(N - K, energy_freq=2, gauss_freq=3)
        fb = ev.export_two_slot(np.float64)
    assert np.array_equal(fa[0], fb[0]) and np.array_equal(fa[1], fb[1]) and fa[2] == fb[2]
    assert len(rows_a) == len(rows_b)
    for a, b in zip(rows_a, rows_b):
        assert a.time_iter == b.time_iter
        assert list(a.energy[:]) == list(b.energy[:]) and list(a.gauss_int[:G]) == list(b.gauss_int[:G])


@pytest.mark.parametrize("G", [5, 7], ids=["g5", "g7"])
def test_family_momentum_generated_state(rt, G):
    """gft_init_grf fills the family links with the identity and the family
    electric field with zero (P = 0, E = 0: the Gauss law holds from the start),
    and the state evolves keeping it (family-symmetric potential)."""
    n = (24, 18, 16)
    d = _desc(G, GFT_FORM_MOMENTUM_F32, n, pot2=FAMILY_SYMMETRIC_POTENTIAL)
    with rt.Evolver(d) as ev:
        ev.init_grf(seed=7, rms=0.1, slope=-1.0, vevs=(0.8, 0.6))
        _, v0, _ = ev.export_two_slot(np.float64)
        rows = ev.evolve(60, energy_freq=30, gauss_freq=30)
    N = int(np.prod(n))
    fam = v0.reshape(2, N, 3, -1)[..., 5:]
    want = np.zeros(fam.shape[-1])
    if G == 7:
        want[0] = 1.0
    assert np.array_equal(fam, np.broadcast_to(want, fam.shape))
    rel = max(r.gauss_int[m] / r.gauss_scale[m] for r in rows if r.has_gauss for m in range(G))
    print(f"gaugeNum {G}: generated state, relative Gauss violation {rel:.2e}")
    assert rel <= 1e-6
    assert all(np.isfinite(r.energy[:]).all() for r in rows if r.has_energy)
