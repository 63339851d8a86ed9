"""Checkpoint / restore and snapshot / rollback (gft_checkpoint, gft_restore,
gft_snapshot, gft_rollback).  The reference writes 6-digit text field dumps
and has no reader (Analysers/Source/OutputFields.cpp:63-133), so resuming is
this library's own and is held to the strictest bar: N steps in one go and
K steps, checkpoint, a fresh context restored from it, N - K steps give
bitwise the same fields and the same measured rows, in every form (the
momentum forms carry both ping-pong buffers, P, E and the residual words, the
reference doublets and the buffer parity).  A domain error rolls the lattice
back to its snapshot bitwise."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from gftevolver_amd._abi import (GFT_EDOMAIN, GFT_EINVAL, GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F32_PLAIN,
                                 GFT_FORM_MOMENTUM_F64, GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64, GFT_MODEL_2HDM,
                                 TWO_HDM_POTENTIAL, make_desc)
from tests import lattice_np as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _obs(rows):
    return np.array([list(r.energy) + list(r.gauss_int) + list(r.gauss_max) + [r.time_iter] for r in rows])


CASES = {
    "two_slot_f64": dict(form=GFT_FORM_TWO_SLOT_F64),
    "two_slot_f32": dict(form=GFT_FORM_TWO_SLOT_F32),
    "momentum_f32": dict(form=GFT_FORM_MOMENTUM_F32),
    "momentum_f32_plain": dict(form=GFT_FORM_MOMENTUM_F32_PLAIN),
    "momentum_f64": dict(form=GFT_FORM_MOMENTUM_F64),
    "momentum_f32_2hdm": dict(form=GFT_FORM_MOMENTUM_F32, model=GFT_MODEL_2HDM, pot2=TWO_HDM_POTENTIAL),
    "momentum_f32_fixed": dict(form=GFT_FORM_MOMENTUM_F32, fixed_faces=1 | 2),
    "two_slot_f32_gen": dict(form=GFT_FORM_TWO_SLOT_F32, gen_rep=1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_resume_is_bitwise(rt, case):
    kw = CASES[case]
    n = (12, 10, 14)
    d = make_desc(n=n, **kw)
    ns = 8 if kw.get("model") == GFT_MODEL_2HDM else 4
    s, v = L.near_vacuum_state(n, ns, seed=3, amp=0.1, mom=0.01, link_eps=0.1)
    if kw.get("fixed_faces"):
        s[0], v[0] = s[1], v[1]     # a frozen shell keeps both slots equal
    if kw.get("gen_rep"):
        # the same links in the generator form: w = atan2(|c|, c0) c / |c|
        vq = v.reshape(2, -1, 3, 5)
        w = np.zeros(vq.shape[:-1] + (4,))
        w[..., 0] = vq[..., 0]
        c = vq[..., 1:]
        nrm = np.linalg.norm(c[..., 1:], axis=-1)
        ang = np.arctan2(nrm, c[..., 0])
        w[..., 1:] = c[..., 1:] * (ang / np.where(nrm > 0, nrm, 1.0))[..., None]
        v = w.reshape(2, -1, 12)
    N, K = 9, 4
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 2)
        rows_a = ev.evolve(N, energy_freq=2, gauss_freq=3)
        fa = ev.export_two_slot(np.float64)
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s, v, 2)
        rows_b = ev.evolve(K, energy_freq=2, gauss_freq=3)
        ck = ev.checkpoint()
    with rt.Evolver(d) as ev:          # a fresh context
        ev.restore(ck)
        rows_b += ev.evolve(N - K, energy_freq=2, gauss_freq=3)
        fb = ev.export_two_slot(np.float64)
        if d.form >= GFT_FORM_MOMENTUM_F32:
            assert ev.observe().has_gauss   # the previous step's buffer came back too
    print(f"{case}: checkpoint of {len(ck)} bytes; {len(rows_a)} measured rows")
    assert fa[2] == fb[2] == 2 + N
    np.testing.assert_array_equal(fa[0], fb[0])
    np.testing.assert_array_equal(fa[1], fb[1])
    np.testing.assert_array_equal(_obs(rows_a), _obs(rows_b))


def test_restore_refuses_other_contexts(rt):
    n = (8, 8, 8)
    s, v = L.near_vacuum_state(n, 4, seed=1)
    with rt.Evolver(make_desc(n=n, form=GFT_FORM_MOMENTUM_F32)) as ev:
        ev.import_two_slot(s, v, 0)
        ck = ev.checkpoint()
    for other in (make_desc(n=(8, 8, 6), form=GFT_FORM_MOMENTUM_F32), make_desc(n=n, form=GFT_FORM_MOMENTUM_F64),
                  make_desc(n=n, form=GFT_FORM_MOMENTUM_F32, fixed_faces=3)):
        with rt.Evolver(other) as ev:
            with pytest.raises(rt.GftError) as e:
                ev.restore(ck)
            assert e.value.code == GFT_EINVAL
    with rt.Evolver(make_desc(n=n, form=GFT_FORM_MOMENTUM_F32)) as ev:
        with pytest.raises(rt.GftError):
            ev.restore(ck[:100])               # truncated
        with pytest.raises(rt.GftError):
            ev.restore(b"x" * len(ck))         # not a checkpoint
        with pytest.raises(rt.GftError):
            ev.checkpoint()                    # no state yet


def test_ensemble_members_resume_bitwise(rt):
    sizes = ((16, 16, 16), (24, 24, 24), (12, 16, 20))
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=nn) for nn in sizes]
    states = [L.near_vacuum_state(nn, 4, seed=10 + i, amp=0.1, mom=0.01, link_eps=0.1) for i, nn in enumerate(sizes)]
    with rt.Ensemble(descs) as ens:
        for m, (s, v) in zip(ens.members, states):
            m.import_two_slot(s, v, 0)
        ens.evolve(9)
        ref = [m.export_two_slot(np.float64) for m in ens.members]
    with rt.Ensemble(descs) as ens:
        for m, (s, v) in zip(ens.members, states):
            m.import_two_slot(s, v, 0)
        ens.evolve(4)
        cks = [m.checkpoint() for m in ens.members]
    with rt.Ensemble(descs) as ens:
        for m, ck in zip(ens.members, cks):
            m.restore(ck)
        ens.evolve(5)
        out = [m.export_two_slot(np.float64) for m in ens.members]
    for a, b in zip(ref, out):
        assert a[2] == b[2] == 9
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("form", [GFT_FORM_MOMENTUM_F32, GFT_FORM_TWO_SLOT_F64], ids=["momentum_f32", "two_slot_f64"])
def test_rollback_after_domain_error(rt, form):
    """dt = 3 on a hot state drives |R| past 1 in a link update: gft_evolve
    reports GFT_EDOMAIN and gft_rollback restores the snapshot bitwise."""
    n = (4, 4, 4)
    d = make_desc(form=form, n=n, dt=3.0)
    s, v = L.near_vacuum_state(n, 4, seed=2, amp=0.5, mom=0.2, link_eps=0.5)
    with rt.Evolver(d) as ev:
        with pytest.raises(rt.GftError):
            ev.rollback()                      # no snapshot yet
        ev.import_two_slot(s, v, 5)
        before = ev.export_two_slot(np.float64)
        ev.snapshot()
        with pytest.raises(rt.GftError) as e:
            ev.evolve(3)
        assert e.value.code == GFT_EDOMAIN
        it = ev.rollback()
        after = ev.export_two_slot(np.float64)
    assert it == before[2] == after[2] == 5
    np.testing.assert_array_equal(before[0], after[0])
    np.testing.assert_array_equal(before[1], after[1])


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "gftevolver_amd", "gft_run")) or
                    not os.path.isdir(os.path.join(ROOT, "oracle", "_ref", "cfg")),
                    reason="gft_run or the reference config template not built")
def test_host_lattice_rolls_back_and_reports(tmp_path):
    """The C++ host Lattice snapshots every chunk: a domain error in the first
    chunk comes back as an error naming the step window and the step the
    lattice was rolled back to."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcfg
    refcfg.prepare(str(tmp_path))
    refcfg.write_model(str(tmp_path), 0, 0.65, 0.35)
    refcfg.write_lattice(str(tmp_path), (6, 5, 7), (0.5, 0.5, 0.5), 12, 3.0, 4, ic_seed=7, ic_range=(-0.8, 0.8),
                         efreq=3, gfreq=2)
    r = subprocess.run([os.path.join(ROOT, "gftevolver_amd", "gft_run"), os.path.join(str(tmp_path), "cfg"),
                        str(tmp_path / "data")], capture_output=True, text=True, timeout=120)
    print(r.stderr[-400:])
    assert r.returncode != 0
    assert "Panic in evolve" in r.stderr and "rolled back to step 0" in r.stderr
