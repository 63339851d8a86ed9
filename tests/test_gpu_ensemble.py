"""GPU checks of the batched ensemble (gft_ensemble_*): independent FP32
momentum-form lattices of different sizes advanced by one kernel launch
per step, with and without a captured two-step hipGraph.

Each member must be bitwise the same lattice evolved alone by the same kernel
(a one-member ensemble with the same tile shape): any difference is a bug in
the work list, the member table or the parity bookkeeping.  The single-lattice
kernel (gft_evolve) evaluates the same expressions (FP contraction is per
expression, -ffp-contract=on), so it must agree bitwise as well, and so must
the ensemble kernel's staging modes (global loads, LDS plane tile, shared
plaquettes)."""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import GFT_FORM_MOMENTUM_F32, make_desc
from tests import lattice_np as L

pytestmark = pytest.mark.gpu

# Ragged sizes: partial z and y tiles, odd extents, one member with nx = 1 row of tiles.
SIZES = [(8, 12, 16), (16, 16, 16), (12, 8, 20), (5, 7, 33), (24, 24, 24)]
TILES = [(64, 4), (16, 16), (32, 8), (256, 1), (128, 2)]


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _states():
    return [L.near_vacuum_state(n, 4, seed=100 + i, amp=0.2, mom=0.02, link_eps=0.1) for i, n in enumerate(SIZES)]


def _run_alone(rt, d, s, v, steps, tile, use_graph):
    with rt.Ensemble([d]) as ens:
        ens.members[0].import_two_slot(s.astype(np.float32), v.astype(np.float32), 0)
        ens.set_launch(0, *tile)
        ens.evolve(steps, use_graph=use_graph)
        return ens.members[0].export_two_slot(np.float32), ens.members[0].observe()


def _run_single_kernel(rt, d, s, v, steps, tile):
    with rt.Evolver(d) as ev:
        ev.import_two_slot(s.astype(np.float32), v.astype(np.float32), 0)
        ev.set_launch(tile[0], tile[1], order=0, min_waves=4)
        ev.evolve(steps)
        return ev.export_two_slot(np.float32)


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "stream"])
@pytest.mark.parametrize("steps", [7, 10])
def test_ensemble_members_bitwise_equal_single_runs(rt, use_graph, steps):
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=n) for n in SIZES]
    states = _states()
    with rt.Ensemble(descs) as ens:
        for i, (m, (s, v)) in enumerate(zip(ens.members, states)):
            m.import_two_slot(s.astype(np.float32), v.astype(np.float32), 0)
            ens.set_launch(i, *TILES[i])
        # Split into two calls: the second starts from the other parity.
        ens.evolve(3, use_graph=use_graph)
        ens.evolve(steps - 3, use_graph=use_graph)
        got = [m.export_two_slot(np.float32) for m in ens.members]
        gobs = [m.observe() for m in ens.members]
    for i, (d, (s, v)) in enumerate(zip(descs, states)):
        # The alone run uses the other launch mode, so graph and stream are compared too.
        (rs, rv, rit), robs = _run_alone(rt, d, s, v, steps, TILES[i], not use_graph)
        gs, gv, git = got[i]
        print(f"member {i} {SIZES[i]} tile {TILES[i]}: time_iter {git}, "
              f"max |diff| scalar {np.abs(gs - rs).max():.3e} links {np.abs(gv - rv).max():.3e}")
        assert git == rit == steps
        assert np.array_equal(gs, rs) and np.array_equal(gv, rv)
        assert gobs[i].energy[0] == robs.energy[0]
        ks, kv, _ = _run_single_kernel(rt, d, s, v, steps, TILES[i])
        ds, dv = np.abs(gs - ks).max(), np.abs(gv - kv).max()
        print(f"  vs gft_evolve: scalar {ds:.3e} links {dv:.3e}")
        assert np.array_equal(gs, ks) and np.array_equal(gv, kv)


def test_ensemble_staging_modes_bitwise_equal(rt):
    """The ensemble kernel with every neighbour from global memory (stage 0),
    with the LDS plane tile (1) and with shared plaquettes (2): the same fields.
    The ensemble runs the lowest stage any member asks for."""
    tiles = [(64, 4), (16, 16), (32, 8), (64, 4), (128, 2)]
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=n) for n in SIZES]
    states = _states()
    out = {}
    for stage in (2, 1, 0):
        with rt.Ensemble(descs) as ens:
            for i, (m, (s, v)) in enumerate(zip(ens.members, states)):
                m.import_two_slot(s.astype(np.float32), v.astype(np.float32), 0)
                m.set_launch(tiles[i][0], tiles[i][1], 2, 4, stage)
            ens.evolve(9)
            out[stage] = [m.export_two_slot(np.float32) for m in ens.members]
    for stage in (1, 0):
        for i in range(len(SIZES)):
            a, b = out[2][i], out[stage][i]
            print(f"stage {stage} member {i}: max |diff| scalar {np.abs(a[0] - b[0]).max():.3e} "
                  f"links {np.abs(a[1] - b[1]).max():.3e}")
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_ensemble_rejects_mixed_steps(rt):
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=n) for n in SIZES[:2]]
    states = _states()[:2]
    with rt.Ensemble(descs) as ens:
        ens.members[0].import_two_slot(states[0][0].astype(np.float32), states[0][1].astype(np.float32), 0)
        with pytest.raises(rt.GftError, match="import every member"):
            ens.evolve(2)
        ens.members[1].import_two_slot(states[1][0].astype(np.float32), states[1][1].astype(np.float32), 5)
        with pytest.raises(rt.GftError, match="same step"):
            ens.evolve(2)


def test_ensemble_workload_observables_match_fp64_sequential(rt):
    """The 64-member ensemble workload (16 x 32^3, 24 x 48^3, 16 x 64^3,
    8 x 96^3), 2000 FP32 steps in one graph-replayed launch per
    step, against the same members run one by one in FP64: the ensemble means
    of <|phi|^2>/v^2, the electric and magnetic energy fractions and the 1x1
    SU(2) and U(1) plaquettes agree to 1e-4 at every measurement (every 500
    steps)."""
    from gftevolver_amd import ensemble as E
    specs = E.members()
    assert len(specs) == 64 and sum(n[0] ** 3 for n, _, _ in specs) == 14450688
    o32, sec, kms, tiles = E.run_fp32(rt, specs, 2000, 500, tune=True, log=print)
    o64 = E.run_fp64_sequential(rt, specs, 2000, 500)
    m32, m64 = o32.mean(axis=1), o64.mean(axis=1)
    for t in range(m32.shape[0]):
        print(f"step {500 * t}: " + ", ".join(f"{k} {a:.7f}/{b:.7f}" for k, a, b in zip(E.OBSERVABLES, m32[t], m64[t])))
    per_member = np.abs(o32 - o64).max(axis=(0, 1))
    print(f"{14450688 * 2000 / sec:.3e} site-updates/s ({kms:.3f} ms per step by HIP events); "
          f"max per-member |diff| {dict(zip(E.OBSERVABLES, per_member))}")
    assert np.abs(m32 - m64).max() <= 1e-4


def test_ensemble_graph_replays_with_eager_work_between(rt):
    """One two-step graph replay per call, with eager work (observables,
    exports) between the replays: after every replay the members equal the
    stream-launched run bitwise."""
    descs = [make_desc(form=GFT_FORM_MOMENTUM_F32, n=n) for n in SIZES[:3]]
    states = _states()[:3]
    with rt.Ensemble(descs) as eg, rt.Ensemble(descs) as es:
        for ens in (eg, es):
            for m, (s, v) in zip(ens.members, states):
                m.import_two_slot(s.astype(np.float32), v.astype(np.float32), 0)
        for rep in range(4):
            eg.evolve(2, use_graph=True)
            es.evolve(2, use_graph=False)
            for a, b in zip(eg.members, es.members):
                a.observe()
                sa, va, ia = a.export_two_slot(np.float32)
                sb, vb, ib = b.export_two_slot(np.float32)
                assert ia == ib == 2 * (rep + 1)
                assert np.array_equal(sa, sb) and np.array_equal(va, vb), f"replay {rep}"
