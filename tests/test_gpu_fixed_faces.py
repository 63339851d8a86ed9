"""GPU parity of fixed faces (the reference's Fixed boundary type, face type 1 in
Core/Config/Lattice.cfg; BoundaryConditions/Source/Fixed.cpp, Core/Source/
Lattice.cpp:287-312): the one-site shell of every fixed face is not updated, is still read
by its neighbours, and is left out of the energy and Gauss-law sums.

The initial states hold both time slots equal on the shell, as the reference
requires for a frozen site (it never writes the shell, so unequal slots would
alternate).  Tolerances are those of the periodic parity tests:
  * two-slot FP64 vs the FP64 CPU oracle: <= 1e-12 relative (fields), energies
    and Gauss integrals <= 1e-10 relative.  With frozen shells the Gauss law is
    violated physically next to the shell (the frozen links carry no current),
    so it is compared, not bounded;
  * momentum form vs the oracle: <= 1e-11 (FP64) / 2e-5 (FP32) relative at
    field level, energies <= 1e-9 / 1e-5 of the total;
  * the shell itself: bitwise unchanged in every form (compared with an
    import/export round trip of the same state);
  * every tile shape, block order and staging mode: bitwise the same fields.
"""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN, GFT_FORM_MOMENTUM_F64,
                                 GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64, GFT_MODEL_2HDM, GFT_MODEL_SM,
                                 TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

# bit 2a: low face of axis a fixed, bit 2a + 1: high face.
FACES = {"xz": 1 | 2 | 16 | 32, "y": 4 | 8, "all": 63}


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _phase_diff(va, vb):
    d = (va - vb).reshape(va.shape[:-1] + (3, 5)).copy()
    d[..., 0] = (d[..., 0] + np.pi) % (2 * np.pi) - np.pi
    return d


def _state(n, faces, nd=1, seed=31):
    """Near-vacuum state with both slots equal on the frozen shell."""
    s, v = L.near_vacuum_state(n, 4 * nd, seed=seed, amp=0.15, mom=0.02, link_eps=0.1,
                               vev=0.7 if nd == 2 else 1.0)
    m = O.frozen_mask(n, faces)
    s[0][m], v[0][m] = s[1][m], v[1][m]
    return s, v, m


def _roundtrip(rt, d, s, v, dtype=np.float64):
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        return ev.export_two_slot(dtype)


@pytest.mark.parametrize("case", list(FACES), ids=list(FACES))
def test_twoslot_f64_fixed_faces_match_oracle(rt, case):
    n = (10, 8, 12)
    faces = FACES[case]
    d = make_desc(form=GFT_FORM_TWO_SLOT_F64, n=n, fixed_faces=faces)
    s, v, m = _state(n, faces)
    s0, v0 = s.copy(), v.copy()
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(30, energy_freq=10, gauss_freq=10)
        gs, gv, it = ev.export_two_slot(np.float64)
    assert it == 30
    _, orows, _ = O.evolve(d, s, v, 30, energy_freq=10, gauss_freq=10, nthreads=8)
    rs, rv = _rel(gs, s), _rel(gv, v)
    print(f"faces {faces}: FP64 two-slot vs oracle after 30 steps: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= 1e-12 and rv <= 1e-12
    assert np.array_equal(gs[:, m], s0[:, m]) and np.array_equal(gv[:, m], v0[:, m])
    assert not np.array_equal(gs[:, ~m], s0[:, ~m])
    assert len(rows) == len(orows) == 3
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter
        np.testing.assert_allclose(np.array(a.energy[:]), np.array(b.energy[:]), rtol=1e-10,
                                   atol=1e-11 * abs(b.energy[0]))
        gi, go = np.array(a.gauss_int[:]), np.array(b.gauss_int[:])
        print(f"  step {a.time_iter}: Gauss integral {gi[0]:.6e} vs {go[0]:.6e}")
        np.testing.assert_allclose(gi, go, rtol=1e-9, atol=1e-11 * max(b.gauss_scale[:]))
        np.testing.assert_allclose(np.array(a.gauss_max[:]), np.array(b.gauss_max[:]), rtol=1e-9,
                                   atol=1e-11 * max(b.gauss_scale[:]))


def test_twoslot_f32_fixed_faces_match_oracle(rt):
    """The reference's own storage (float fields, FP64 arithmetic): a few float
    ulps from the FP64 oracle run on the same float arrays."""
    n = (6, 5, 7)
    faces = FACES["xz"]
    d = make_desc(form=GFT_FORM_TWO_SLOT_F32, n=n, fixed_faces=faces)
    s, v, m = _state(n, faces)
    s, v = s.astype(np.float32), v.astype(np.float32)
    s0, v0 = s.copy(), v.copy()
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        ev.evolve(12)
        gs, gv, _ = ev.export_two_slot(np.float32)
    O.evolve(d, s, v, 12, nthreads=4)
    ds, dv = np.abs(gs - s).max(), np.abs(gv - v).max()
    print(f"FP32 two-slot vs float oracle: max |diff| scalar {ds:.2e} links {dv:.2e}")
    assert ds <= 2e-6 and dv <= 2e-6
    assert np.array_equal(gs[:, m], s0[:, m]) and np.array_equal(gv[:, m], v0[:, m])


@pytest.mark.parametrize("case", ["xz", "all"])
@pytest.mark.parametrize("form,tol", [(GFT_FORM_MOMENTUM_F64, 1e-11), (GFT_FORM_MOMENTUM_F32, 2e-5),
                                      (GFT_FORM_MOMENTUM_F32_PLAIN, 2e-5)], ids=["f64", "f32", "f32plain"])
def test_momentum_fixed_faces_match_oracle(rt, form, tol, case):
    n = (16, 12, 20)
    faces = FACES[case]
    d = make_desc(form=form, n=n, fixed_faces=faces)
    s, v, m = _state(n, faces)
    base_s, base_v, _ = _roundtrip(rt, d, s, v)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 0)
        rows = ev.evolve(60, energy_freq=20, gauss_freq=20)
        gs, gv, _ = ev.export_two_slot(np.float64)
    _, orows, _ = O.evolve(d, s, v, 60, energy_freq=20, gauss_freq=20, nthreads=8)
    cur = (60 + 1) % 2
    rs = _rel(gs[cur], s[cur])
    rv = np.abs(_phase_diff(gv[cur], v[cur])).max()
    print(f"form {form} faces {faces}: fields vs FP64 oracle after 60 steps: scalar {rs:.2e}, links {rv:.2e}")
    assert rs <= tol and rv <= tol
    # The shell is carried unchanged through every step (both slots).
    assert np.array_equal(gs[:, m], base_s[:, m]) and np.array_equal(gv[:, m], base_v[:, m])
    etol = 1e-9 if form == GFT_FORM_MOMENTUM_F64 else 1e-5
    for a, b in zip(rows, orows):
        assert a.time_iter == b.time_iter
        err = np.abs(np.array(a.energy[:]) - np.array(b.energy[:])).max() / abs(b.energy[0])
        gi, go = np.array(a.gauss_int[:]), np.array(b.gauss_int[:])
        gerr = np.abs(gi - go).max() / max(np.abs(go).max(), 1e-300)
        print(f"  step {a.time_iter}: energy rel col err {err:.1e}; Gauss {gi[0]:.4e} vs {go[0]:.4e} ({gerr:.1e})")
        assert err <= etol
        # The shell-induced violation is O(field change), far above rounding.
        assert gerr <= (1e-8 if form == GFT_FORM_MOMENTUM_F64 else 1e-3)


def test_momentum_2hdm_fixed_faces_match_oracle(rt):
    n = (8, 6, 10)
    faces = FACES["y"] | 1 | 2
    for form, tol in ((GFT_FORM_MOMENTUM_F64, 1e-11), (GFT_FORM_MOMENTUM_F32, 2e-5)):
        d = make_desc(model=GFT_MODEL_2HDM, form=form, n=n, pot2=TWO_HDM_POTENTIAL, fixed_faces=faces)
        s, v, m = _state(n, faces, nd=2, seed=9)
        base_s, base_v, _ = _roundtrip(rt, d, s, v)
        with rt.Evolver(d) as ev:
            ev.import_two_slot(s, v, 0)
            ev.evolve(40)
            gs, gv, _ = ev.export_two_slot(np.float64)
        O.evolve(d, s, v, 40, nthreads=8)
        cur = (40 + 1) % 2
        rs, rv = _rel(gs[cur], s[cur]), np.abs(_phase_diff(gv[cur], v[cur])).max()
        print(f"2HDM form {form}: scalar {rs:.2e} links {rv:.2e}")
        assert rs <= tol and rv <= tol
        assert np.array_equal(gs[:, m], base_s[:, m]) and np.array_equal(gv[:, m], base_v[:, m])


@pytest.mark.parametrize("case", list(FACES), ids=list(FACES))
def test_fixed_faces_launch_shapes_bitwise_identical(rt, case):
    """The frozen check sits before any neighbour gather in the global-load and
    plane-tile kernels and after the shared-plaquette exchange in the stage-2
    kernel (the shell's plaquettes serve its neighbours): every tile shape,
    block order and staging mode must give the same fields."""
    n = (16, 12, 20)
    faces = FACES[case]
    d = make_desc(form=GFT_FORM_MOMENTUM_F32, n=n, fixed_faces=faces)
    s, v, _ = _state(n, faces, seed=6)
    cfgs = ((64, 4, 2, 4, 2), (64, 4, 2, 4, 1), (64, 4, 2, 4, 0), (64, 4, 0, 4, 2), (256, 1, 2, 4, 2),
            (128, 2, 1, 4, 2), (32, 8, 2, 4, 2), (32, 8, 2, 4, 1), (16, 16, 2, 4, 2), (16, 16, 0, 4, 0),
            (64, 4, 3, 4, 2), (32, 8, 4, 4, 1))
    out = {}
    for c in cfgs:
        with rt.Evolver(d) as ev:
            ev.set_launch(*c)
            ev.import_two_slot(s, v, 0)
            ev.evolve(7)
            out[c] = ev.export_two_slot(np.float64)
    ref = out[cfgs[0]]
    bad = []
    for c in cfgs[1:]:
        nd_ = int((out[c][0] != ref[0]).sum() + (out[c][1] != ref[1]).sum())
        print(f"faces {faces} launch {c}: {nd_} differing values")
        if nd_:
            bad.append(c)
    assert not bad, bad


def test_ensemble_members_with_fixed_faces_bitwise_equal_single_runs(rt):
    """Members with different face masks in one ensemble launch: each equals
    the same lattice evolved alone by gft_evolve."""
    sizes = [(8, 12, 16), (12, 8, 20), (16, 16, 16)]
    faces = [FACES["xz"], FACES["y"], 0]
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=n, fixed_faces=f) for n, f in zip(sizes, faces)]
    states = [_state(n, f, seed=50 + i)[:2] for i, (n, f) in enumerate(zip(sizes, faces))]
    states = [(s.astype(np.float32), v.astype(np.float32)) for s, v in states]
    with rt.Ensemble(descs) as ens:
        for i, (mbr, (s, v)) in enumerate(zip(ens.members, states)):
            mbr.import_two_slot(s, v, 0)
            mbr.set_launch(64, 4, 0, 3, 2)
        ens.evolve(9)
        got = [mbr.export_two_slot(np.float32) for mbr in ens.members]
    for i, (d, (s, v)) in enumerate(zip(descs, states)):
        with rt.Evolver(d) as ev:
            ev.import_two_slot(s, v, 0)
            ev.set_launch(64, 4, 0, 3, 2)
            ev.evolve(9)
            ks, kv, _ = ev.export_two_slot(np.float32)
        print(f"member {i} faces {faces[i]}: max |diff| scalar {np.abs(got[i][0] - ks).max():.3e} "
              f"links {np.abs(got[i][1] - kv).max():.3e}")
        assert np.array_equal(got[i][0], ks) and np.array_equal(got[i][1], kv)


@pytest.mark.parametrize("faces", [1, 2 | 16 | 32, 64, -1])
def test_unpaired_or_invalid_faces_rejected(rt, faces):
    """A periodic face must face a periodic face (Core/Source/Lattice.cpp:244-259); masks
    outside 0..63 are invalid."""
    d = make_desc(form=GFT_FORM_MOMENTUM_F32, n=(8, 8, 8), fixed_faces=faces)
    with pytest.raises(rt.GftError) as e:
        rt.Evolver(d)
    print(e.value)
    assert e.value.code < 0
    if 0 <= faces <= 63:
        assert "both faces of an axis" in str(e.value)
