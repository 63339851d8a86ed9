"""GPU parity of the momentum-form (structure-of-arrays kick/drift) evolver.

The momentum form reorders the reference's arithmetic (stored E instead of
vec(U(t) U(t-dt)^dag), staples shared between orientations, no intermediate
renormalisations), so parity is to a tolerance, not bitwise:
  * FP64 storage vs the FP64 CPU oracle: <= 1e-11 relative at field level after
    100 steps of the 32^3 SM workload (reordering changes the last bits only);
  * FP32 storage vs the FP64 CPU oracle: <= 2e-5 relative at field level after
    100 steps (float rounding of ~1e-7 per step, accumulated);
  * energies: <= 1e-9 (FP64) / 1e-5 (FP32) relative to the total energy;
  * Gauss law, relative to the magnitude of the cancelling terms
    (gft_obs.gauss_scale): <= 1e-12 (FP64) and <= 1e-6 (FP32) over 200 steps
    of a hot state from an exactly constraint-satisfying start.
Every tile shape and block order must give bitwise identical fields (the autotuner
relies on it); the occupancy variants are separate kernels, equal to 2 ulps.
"""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN, GFT_FORM_MOMENTUM_F64,
                                 GFT_FORM_TWO_SLOT_F64,
                                 GFT_MODEL_2HDM, GFT_MODEL_SM, TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L
from tests import oracle_lib as O
from tests.test_gpu_twoslot import gaussian_ic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _phase_diff(va, vb):
    """Link arrays compared with phases modulo 2 pi (the momentum form wraps them)."""
    d = (va - vb).reshape(va.shape[:-1] + (3, 5)).copy()
    d[..., 0] = (d[..., 0] + np.pi) % (2 * np.pi) - np.pi
    return d


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-11), (GFT_FORM_MOMENTUM_F32, 2e-5),
                                      (GFT_FORM_MOMENTUM_F32_PLAIN, 2e-5)], ids=["f64", "f32", "f32plain"])
def test_momentum_matches_oracle_32cube_100_steps(rt, form, tol):
    n = (32, 32, 32)
    d = make_desc(form=form, n=n)
    s, v = gaussian_ic(n)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(100, energy_freq=20)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 100
    _, orows, _ = O.evolve(d, s, v, 100, energy_freq=20, nthreads=8)
    cur = (100 + 1) % 2
    rs = _rel(gs[cur], s[cur])
    rv = np.abs(_phase_diff(gv[cur], v[cur])).max()
    print(f"form {form}: field error vs FP64 oracle after 100 steps: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= tol and rv <= tol
    etol = 1e-9 if form == GFT_FORM_MOMENTUM_F64 else 1e-5
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter
        err = np.abs(np.array(a.energy[:]) - np.array(b.energy[:])).max() / abs(b.energy[0])
        print(f"  step {a.time_iter}: energy {a.energy[0]:.9e} vs {b.energy[0]:.9e} (rel col err {err:.1e})")
        assert err <= etol


@pytest.mark.parametrize("form", [GFT_FORM_MOMENTUM_F64, GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN],
                         ids=["f64", "f32", "f32plain"])
def test_momentum_2hdm_matches_oracle(rt, form):
    n = (8, 6, 10)
    d = make_desc(model=GFT_MODEL_2HDM, form=form, n=n, pot2=TWO_HDM_POTENTIAL)
    s, v = L.near_vacuum_state(n, 8, seed=9, amp=0.1, mom=0.01, link_eps=0.05, vev=0.7)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        ev.evolve(40)
        gs, gv, _ = ev.export_two_slot(np.float64)
    O.evolve(d, s, v, 40, nthreads=8)
    cur = (40 + 1) % 2
    tol = 1e-11 if form == GFT_FORM_MOMENTUM_F64 else 2e-5
    assert _rel(gs[cur], s[cur]) <= tol
    assert np.abs(_phase_diff(gv[cur], v[cur])).max() <= tol


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-12), (GFT_FORM_MOMENTUM_F32, 1e-6),
                                      (GFT_FORM_MOMENTUM_F32_PLAIN, 1e-6)], ids=["f64", "f32", "f32plain"])
def test_momentum_gauss_law_conserved(rt, form, tol):
    """From E = 0, P = 0 the discrete Gauss law holds exactly; it must stay at
    rounding level.  Ragged lattice, unequal spacings, non-trivial links: a hot
    state whose plaquettes are far from 1, so the FP32 rounding of the forces
    themselves (not of the storage) sets the FP32 level here."""
    n = (12, 10, 14)
    d = make_desc(form=form, n=n, d=(0.5, 0.45, 0.55))
    s, v = L.near_vacuum_state(n, 4, seed=4, amp=0.2, mom=0.0, link_eps=0.2)
    s[0], v[0] = s[1], v[1]
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(200, gauss_freq=50)
    rel = [max(r.gauss_int[m] / r.gauss_scale[m] for m in range(4)) for r in rows]
    print(f"form {form}: relative Gauss violation every 50 steps: {['%.2e' % x for x in rel]}")
    assert max(rel) <= tol


@pytest.mark.parametrize("form,model", [(GFT_FORM_MOMENTUM_F32, GFT_MODEL_SM), (GFT_FORM_MOMENTUM_F32_PLAIN, GFT_MODEL_SM),
                                        (GFT_FORM_MOMENTUM_F32, GFT_MODEL_2HDM)], ids=["f32", "f32plain", "f32-2hdm"])
def test_momentum_launch_shapes_bitwise_identical(rt, form, model):
    """Tile shape, block order, LDS staging of the in-plane neighbours and the
    sharing of in-plane plaquettes and covariant differences change where
    values come from, never the arithmetic: bitwise the same fields.  The side
    lengths (16, 12, 20) are shorter than some tiles (bz = 256 > nz, by = 16 >
    ny), so the staged ring wraps more than once and ragged tiles leave threads
    without a site; (256, 1) runs stage 2 as stage 1 (bz + by > 256)."""
    n = (16, 12, 20)
    nd = 2 if model == GFT_MODEL_2HDM else 1
    d = make_desc(model=model, form=form, n=n, pot2=TWO_HDM_POTENTIAL if nd == 2 else None)
    s, v = L.near_vacuum_state(n, 4 * nd, seed=6, amp=0.1, mom=0.01, link_eps=0.05)
    # (bz, by, order, min_waves, stage); min_waves 4 is the default kernel
    cfgs = ((64, 4, 2, 4, 2), (64, 4, 2, 4, 1), (64, 4, 2, 4, 0), (64, 4, 0, 4, 2), (64, 4, 1, 4, 1),
            (256, 1, 2, 4, 2), (256, 1, 2, 4, 0), (128, 2, 2, 4, 2), (128, 2, 2, 4, 1), (32, 8, 1, 4, 2),
            (32, 8, 1, 4, 0), (16, 16, 0, 4, 2), (16, 16, 2, 4, 1), (16, 16, 2, 4, 0), (64, 4, 2, 0, 2),
            (64, 4, 2, 0, 1), (64, 4, 2, 0, 0), (64, 4, 2, 3, 2), (64, 4, 2, 3, 1), (64, 4, 2, 3, 0),
            (64, 4, 3, 4, 2), (32, 8, 4, 4, 2), (16, 16, 3, 4, 1), (128, 2, 4, 4, 0))
    out = {}
    for c in cfgs:
        with rt.Evolver(d) as ev:
            ev.set_launch(*c)
            ev.import_two_slot(s, v, 0)
            ev.evolve(7)
            out[c] = ev.export_two_slot(np.float64)
    bad = []
    for c in cfgs[1:]:
        # The 2HDM kernel has one occupancy variant; the SM kernel's min_waves
        # variants are separately compiled, compared within their own variant.
        ref_c = cfgs[0] if (nd == 2 or c[3] == 4) else (64, 4, 2, c[3], 2)
        if ref_c == c:
            continue
        ref = out[ref_c]
        ds = np.abs(out[c][0] - ref[0]).max()
        dv = np.abs(out[c][1] - ref[1]).max()
        nd_ = int((out[c][0] != ref[0]).sum() + (out[c][1] != ref[1]).sum())
        print(f"launch {c} vs {ref_c}: max |diff| scalar {ds:.2e} links {dv:.2e}, {nd_} differing values")
        if nd_:
            bad.append(c)
    if nd == 1:
        # Across occupancy variants the register allocation differs, the
        # arithmetic does not: at most two float ulps at the field's scale.
        ref = out[cfgs[0]]
        for mw in (0, 3):
            o = out[(64, 4, 2, mw, 2)]
            ds = np.abs(o[0] - ref[0]).max()
            dv = np.abs(o[1] - ref[1]).max()
            print(f"min_waves {mw} vs 4: max |diff| scalar {ds:.2e} links {dv:.2e}")
            if ds > 2 * np.spacing(np.float32(np.abs(ref[0]).max())) or dv > 2 * np.spacing(np.float32(1)):
                bad.append(("min_waves", mw))
    assert not bad, bad


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F32, 8e-9), (GFT_FORM_MOMENTUM_F32_PLAIN, 1.2e-7)],
                         ids=["f32", "f32plain"])
def test_momentum_f32_storage_precision(rt, form, tol):
    """An FP64 state survives import and export to the storage's precision.
    phi is stored as the FP32 difference from a reference doublet (the mean of
    the imported phi, rounded to FP32): its error is at most half an ulp of the
    fluctuation phi - ref, i.e. <= 2^-24 of max |phi - ref|.  P (and E): FP32 +
    4-bit residual (q ulp(hi) / 2^3, error at most 2^-4 ulp, <= 2^-27 of the
    largest value), plain FP32 24 bits.  P is phi(t) - phi(t-dt), so the
    previous slot carries P's precision."""
    n = (6, 5, 4)
    d = make_desc(form=form, n=n)
    s, v = L.near_vacuum_state(n, 4, seed=7, amp=0.2, mom=0.02, link_eps=0.2)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        gs, gv, _ = ev.export_two_slot(np.float64)
    cur, old = 1, 0
    ref = s[cur].mean(axis=0).astype(np.float32).astype(np.float64)
    e_phi = np.abs(gs[cur] - s[cur]).max() / np.abs(s[cur] - ref).max()
    e_p = np.abs((gs[cur] - gs[old]) - (s[cur] - s[old])).max() / np.abs(s[cur] - s[old]).max()
    print(f"form {form}: error of phi over max |phi - ref| {e_phi:.2e}, relative error of P {e_p:.2e}")
    assert e_phi <= 2.0 ** -24 * 1.0001 and e_p <= tol


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F32, 6e-7), (GFT_FORM_MOMENTUM_F32_PLAIN, 2.5e-6)],
                         ids=["f32", "f32plain"])
def test_momentum_gauss_law_after_1e4_steps(rt, form, tol):
    """SM 64^3 from a GRF Higgs state (P(k) ~ 1/k, rms 0.1 v, links identity,
    E = 0) for 10^4 steps: the relative Gauss violation (integrated |G| over the
    magnitude of the cancelling terms, per generator) stays near 3e-7 with
    compensated storage (the random walk of the FP32 force rounding) and
    reaches ~1.2e-6 with plain FP32 storage (the random walk of its storage
    rounding)."""
    from gftevolver_amd.ic import grf_state
    n = (64, 64, 64)
    s, v = grf_state(n, rms=0.1, seed=20261203, dtype=np.float32)
    with rt.Evolver(make_desc(form=form, n=n)) as ev:
        ev.import_two_slot(s, v, 0)
        rel = []
        for _ in range(4):
            ev.evolve(2500)
            o = ev.observe()
            rel.append(max(o.gauss_int[m] / o.gauss_scale[m] for m in range(4)))
    print(f"form {form}: relative Gauss violation every 2500 steps: {['%.2e' % x for x in rel]}")
    assert max(rel) <= tol


@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-11), (GFT_FORM_MOMENTUM_F32, 2e-5),
                                      (GFT_FORM_MOMENTUM_F32_PLAIN, 2e-5)], ids=["f64", "f32", "f32plain"])
def test_momentum_damping_and_ragged_lattice(rt, form, tol):
    """Damped steps (Model::update: the first nt_damped steps of the reference's
    step counter scale P and E by 1 - gamma dt / 2) on a lattice with unequal
    sides and spacings, starting mid-schedule (time_iter 3), against the oracle."""
    n = (5, 7, 3)
    d = make_desc(form=form, n=n, d=(0.5, 0.4, 0.6), damping=1.0, nt_damped=5)
    s, v = L.near_vacuum_state(n, 4, seed=2, amp=0.2, mom=0.02, link_eps=0.1)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 3)
        ev.evolve(6)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 9
    O.evolve(d, s, v, 6, time_iter=3)
    cur = (9 + 1) % 2
    rs = _rel(gs[cur], s[cur])
    rv = np.abs(_phase_diff(gv[cur], v[cur])).max()
    print(f"form {form}: damped run vs oracle: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= tol and rv <= tol


def test_momentum_import_export_roundtrip(rt):
    n = (6, 5, 4)
    d = make_desc(form=GFT_FORM_MOMENTUM_F64, n=n)
    s, v = L.near_vacuum_state(n, 4, seed=7, amp=0.2, mom=0.02, link_eps=0.2)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 5)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 5
    assert np.abs(gs - s).max() < 1e-14
    assert np.abs(_phase_diff(gv, v)).max() < 1e-13


def test_momentum_energy_matches_two_slot_reference_definition(rt):
    """gft_observe's energy equals the reference's mixed-time energy of the same
    state (as measured by the FP64 oracle before a step)."""
    n = (8, 8, 8)
    d = make_desc(form=GFT_FORM_MOMENTUM_F64, n=n)
    s, v = L.near_vacuum_state(n, 4, seed=8, amp=0.1, mom=0.01, link_eps=0.05)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        o = ev.observe()
    _, orows, _ = O.evolve(d, s.copy(), v.copy(), 1, energy_freq=1)
    np.testing.assert_allclose(np.array(o.energy[:]), np.array(orows[0].energy[:]), rtol=1e-9,
                               atol=1e-11 * abs(orows[0].energy[0]))


def test_wilson_loops_identity_and_plaquette_energy_relation(rt):
    """Identity links give W = 1; the mean 1x1 loops reproduce the magnetic
    energy: E_mag = N dV sum_{i<j} 4/(dx_i^2 dx_j^2) [(1 - W_su2)/g^2 + (1 - W_u1)/g'^2]
    for an isotropic lattice."""
    n = (6, 6, 6)
    d = make_desc(form=GFT_FORM_MOMENTUM_F64, n=n)
    s, v = L.near_vacuum_state(n, 4, seed=8, amp=0.1, mom=0.0, link_eps=0.0)
    v = O.identity_links(216, np.float64)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        w = ev.wilson_loops()
    assert abs(w.w11_su2 - 1) < 1e-14 and abs(w.w22_su2 - 1) < 1e-14 and abs(w.w11_u1 - 1) < 1e-14
    assert abs(w.phi_sq - (s[1] ** 2).sum(axis=1).mean()) < 1e-12
    s, v = L.near_vacuum_state(n, 4, seed=9, amp=0.1, mom=0.0, link_eps=0.3)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        w = ev.wilson_loops()
        o = ev.observe()
    dx = d.dx
    N = 216
    emag = N * dx ** 3 * 3 * 4 / dx ** 4 * ((1 - w.w11_su2) / d.g ** 2 + (1 - w.w11_u1) / d.gp ** 2)
    assert abs(emag - o.energy[4]) <= 1e-10 * abs(o.energy[4])
    assert w.w22_su2 < w.w11_su2 < 1


def test_momentum_domain_error_is_reported(rt):
    n = (4, 4, 4)
    d = make_desc(form=GFT_FORM_MOMENTUM_F32, n=n, dt=3.0)
    s, v = L.near_vacuum_state(n, 4, seed=2, amp=0.5, mom=0.2, link_eps=0.5)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        with pytest.raises(rt.GftError) as e:
            ev.evolve(3)
        assert e.value.code == -4


def test_bench_streams_measures_and_guards(rt):
    """gft_bench_streams: a positive duration for the FP32 SM forms, refused for
    FP64, and the overwritten state must be generated again before an ensemble
    evolves it."""
    from gftevolver_amd._abi import GFT_FORM_MOMENTUM_F64
    from gftevolver_amd.runtime import GftError
    for form in (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN):
        with rt.Evolver(make_desc(form=form, n=(32, 32, 32))) as ev:
            ev.init_grf(seed=3)
            ms = ev.bench_streams(3)
            print(f"form {form}: stream floor {ms * 1e3:.1f} us for 32^3")
            assert 0.0 < ms < 100.0
    with rt.Evolver(make_desc(form=GFT_FORM_MOMENTUM_F64, n=(8, 8, 8))) as ev:
        with pytest.raises(GftError):
            ev.bench_streams(1)
    with rt.Ensemble([make_desc(form=GFT_FORM_MOMENTUM_F32, n=(16, 16, 16))]) as ens:
        m = ens.members[0]
        m.init_grf(seed=3)
        m.bench_streams(1)
        with pytest.raises(GftError):
            ens.evolve(1)
        m.init_grf(seed=3)
        ens.evolve(1)
