"""Pins the CPU oracle before it is trusted as the checker of the HIP evolver.

1. Golden vectors written by the reference's own compiled code
   (tests/golden/make_golden.py): fields, per-site energy densities and
   per-site Gauss violations must be bitwise equal; the full program's printed
   global_quantities files must be reproduced digit for digit.
2. What the mathematics fixes, for the FP64 path and the FP64 diagnostics that
   no golden covers: gauge covariance of the step and gauge invariance of the
   energy and Gauss law, exact time reversibility, the free-field dispersion
   relation of the leapfrog, force = -dH/d(field) by central differences, and
   Gauss-law conservation from a constraint-satisfying state.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from gftevolver_amd._abi import GFT_MODEL_2HDM, GFT_MODEL_SM, TWO_HDM_POTENTIAL, make_desc
from tests import lattice_np as L
from tests import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODELS = [(GFT_MODEL_SM, "sm", None), (GFT_MODEL_2HDM, "2hdm", TWO_HDM_POTENTIAL)]


@pytest.mark.parametrize("model,tag,pot2", MODELS, ids=["sm", "2hdm"])
def test_oracle_bitwise_vs_reference_harness(model, tag, pot2):
    g = np.load(os.path.join(GOLD, f"harness_{tag}.npz"))
    s, v = g["scalar0"].copy(), g["vec0"].copy()
    n = (4, 5, 6)
    d = make_desc(model=model, n=n, pot2=pot2)
    for k in range(6):
        if k % 3 == 0:
            dens = O.energy_density_f32(d, s, v, k)
            assert np.array_equal(dens, g[f"dens_{k}"]), f"energy densities differ at step {k}"
        O.evolve(d, s, v, 1, time_iter=k, energy_freq=3, gauss_freq=3)
        if k % 3 == 0:
            viol = O.gauss_site_f32(d, s, v, k)
            assert np.array_equal(viol, g[f"viol_{k}"]), f"Gauss violations differ at step {k}"
        if f"scalar_{k}" in g:
            assert np.array_equal(s, g[f"scalar_{k}"]), f"scalar fields differ after step {k}"
            assert np.array_equal(v, g[f"vec_{k}"]), f"link fields differ after step {k}"


@pytest.mark.parametrize("model,tag,pot2", MODELS, ids=["sm", "2hdm"])
def test_oracle_reproduces_reference_program_output(model, tag, pot2):
    g = np.load(os.path.join(GOLD, f"program_{tag}.npz"))
    n, Ns = (6, 5, 7), 4 if model == GFT_MODEL_SM else 8
    N = int(np.prod(n))
    s = O.random_uniform(2 * N * Ns, 7, -0.8, 0.8).reshape(2, N, Ns)
    v = O.identity_links(N)
    d = make_desc(model=model, n=n, pot2=pot2)
    ref, obs, _ = O.evolve(d, s, v, 12, energy_freq=3, gauss_freq=2)
    fmt = np.vectorize(lambda x: float("%.6g" % x))
    e = np.array([r[:6] for r, o in zip(ref, obs) if o.has_energy])
    gs = np.array([r[6:] for r, o in zip(ref, obs) if o.has_gauss])
    assert np.array_equal(fmt(e), g["energy"])
    assert np.array_equal(fmt(gs), g["gauss"])
    # FP64 diagnostics of the same states agree with the float ones to float accuracy.
    e64 = np.array([o.energy[:] for o in obs if o.has_energy])
    np.testing.assert_allclose(e64, e, rtol=2e-5, atol=1e-3 * np.abs(e).max())


@pytest.mark.parametrize("model,tag,pot2", MODELS, ids=["sm", "2hdm"])
def test_oracle_generator_form_bitwise_vs_reference_harness(model, tag, pot2):
    """Links stored as generator components w (U = exp(i w.sigma)): the
    reference converts every read to a quaternion and every updated link back
    (WilsonLoops/Source/StandardModel.cpp:78-157, 456-514).  The fixture's state
    has |w| up to 3 (the pi - asin branch of the inverse) and zero links (the
    tolerance branch)."""
    g = np.load(os.path.join(GOLD, f"harness_{tag}_gen.npz"))
    s, v = g["scalar0"].copy(), g["vec0"].copy()
    assert v.shape[-1] == 12
    w = np.linalg.norm(v.reshape(2, -1, 3, 4)[..., 1:], axis=-1)
    assert (w > np.pi / 2).any() and (w == 0).any()
    d = make_desc(model=model, n=(4, 5, 6), pot2=pot2, gen_rep=1)
    for k in range(6):
        if k % 3 == 0:
            assert np.array_equal(O.energy_density_f32(d, s, v, k), g[f"dens_{k}"]), f"energy densities, step {k}"
        O.evolve(d, s, v, 1, time_iter=k, energy_freq=3, gauss_freq=3)
        if k % 3 == 0:
            assert np.array_equal(O.gauss_site_f32(d, s, v, k), g[f"viol_{k}"]), f"Gauss violations, step {k}"
        if f"scalar_{k}" in g:
            assert np.array_equal(s, g[f"scalar_{k}"]), f"scalar fields differ after step {k}"
            assert np.array_equal(v, g[f"vec_{k}"]), f"link fields differ after step {k}"


@pytest.mark.parametrize("model,tag,pot2", MODELS, ids=["sm", "2hdm"])
def test_oracle_generator_form_reproduces_reference_program_output(model, tag, pot2):
    """The unmodified reference program with 4 vector components per direction
    (and TwoHDM.cfg's generator flag for the 2HDM): printed energies and Gauss
    violations digit for digit.  Its Gauss violations differ from the quaternion
    run's in the printed digits, so the form is exercised."""
    g = np.load(os.path.join(GOLD, f"program_{tag}_gen.npz"))
    n, Ns = (6, 5, 7), 4 if model == GFT_MODEL_SM else 8
    N = int(np.prod(n))
    s = O.random_uniform(2 * N * Ns, 7, -0.8, 0.8).reshape(2, N, Ns)
    v = np.zeros((2, N, 12), np.float32)
    d = make_desc(model=model, n=n, pot2=pot2, gen_rep=1)
    ref, obs, _ = O.evolve(d, s, v, 12, energy_freq=3, gauss_freq=2)
    fmt = np.vectorize(lambda x: float("%.6g" % x))
    e = np.array([r[:6] for r, o in zip(ref, obs) if o.has_energy])
    gs = np.array([r[6:] for r, o in zip(ref, obs) if o.has_gauss])
    assert np.array_equal(fmt(e), g["energy"])
    assert np.array_equal(fmt(gs), g["gauss"])
    assert not np.array_equal(g["gauss"], np.load(os.path.join(GOLD, f"program_{tag}.npz"))["gauss"])


def test_generator_form_conversions_round_trip_and_match_closed_forms():
    """(1) Only x links non-trivial, each a function of x alone, identical slots,
    zero scalar field: every plaquette is 1, there is no
    current and E = 0, so the update
    multiplies by the identity and the stored w must come back from
    exp(i w.sigma) -> w to rounding, for |w| on both sides of pi/2 (the
    pi - asin branch).  (2) The quaternion and generator forms of the same
    state evolve to the same group elements (w -> cos|w| + i sin|w| w/|w|.sigma,
    the textbook exponential map)."""
    n = (3, 3, 3)
    N = 27
    rng = np.random.default_rng(5)
    wx = rng.normal(size=(3, 3))
    wx *= np.array([[0.4], [2.0], [2.9]]) / np.linalg.norm(wx, axis=-1, keepdims=True)
    w = np.repeat(wx, 9, axis=0)          # site index (x * 3 + y) * 3 + z
    assert (np.linalg.norm(w, axis=-1) > np.pi / 2).any()
    vg = np.zeros((2, N, 3, 4))
    vg[:, :, 0, 1:] = w
    s = np.zeros((2, N, 4))
    dg = make_desc(n=n, gen_rep=1)
    sg, vgg = s.copy(), vg.reshape(2, N, 12).copy()
    O.evolve(dg, sg, vgg, 1)
    np.testing.assert_allclose(vgg.reshape(2, N, 3, 4)[0, :, 0, 1:], w, rtol=0, atol=1e-13)
    assert np.array_equal(vgg.reshape(2, N, 3, 4)[0, :, 1:], vg[0, :, 1:])
    w = rng.normal(size=(N, 3, 3))
    w *= rng.uniform(0.1, 3.0, size=(N, 3, 1)) / np.linalg.norm(w, axis=-1, keepdims=True)
    vg = np.zeros((2, N, 3, 4))
    vg[..., 1:] = w
    # Same state in the quaternion form: c0 = cos|w|, c = sin|w| w/|w|.
    ang = np.linalg.norm(w, axis=-1, keepdims=True)
    q = np.concatenate([np.cos(ang), np.sin(ang) * w / ang], axis=-1)
    vq = np.zeros((2, N, 3, 5))
    vq[..., 1:] = q
    rng2 = np.random.default_rng(6)
    s2 = rng2.normal(size=(2, N, 4)) * 0.3
    s2[1] = s2[0]
    dq = make_desc(n=n, g=0.65, gp=0.35)
    dg = make_desc(n=n, gen_rep=1, g=0.65, gp=0.35)
    sq_, vq_ = s2.copy(), vq.reshape(2, N, 15).copy()
    sg_, vg_ = s2.copy(), vg.reshape(2, N, 12).copy()
    O.evolve(dq, sq_, vq_, 3)
    O.evolve(dg, sg_, vg_, 3)
    np.testing.assert_allclose(sg_, sq_, rtol=0, atol=1e-12)
    wg = vg_.reshape(2, N, 3, 4)[..., 1:]
    a2 = np.linalg.norm(wg, axis=-1, keepdims=True)
    qg = np.concatenate([np.cos(a2), np.sin(a2) * wg / a2], axis=-1)
    np.testing.assert_allclose(qg, vq_.reshape(2, N, 3, 5)[..., 1:], rtol=0, atol=1e-12)


FAMILY = [(5, False), (5, True), (7, False), (7, True)]
FAMILY_IDS = ["g5", "g5-gen", "g7", "g7-gen"]


def identity_links_family(N, gauge_num, gen):
    """Zero gauge fields as the reference initialises them (Lattice.cpp:369-392):
    c0 = 1 (and c0' = 1 for gaugeNum 7) in the quaternion form, zeros otherwise."""
    from gftevolver_amd._abi import link_reals
    from types import SimpleNamespace
    nvd = link_reals(SimpleNamespace(gen_rep=int(gen), gauge_num=gauge_num)) // 3
    v = np.zeros((2, N, 3, nvd), np.float32)
    if not gen:
        v[..., 1] = 1.0
        if gauge_num == 7:
            v[..., 5] = 1.0
    return v.reshape(2, N, 3 * nvd)


@pytest.mark.parametrize("gauge_num,gen", FAMILY, ids=FAMILY_IDS)
def test_oracle_family_gauge_bitwise_vs_reference_harness(gauge_num, gen):
    """The 2HDM with a Higgs-family U(1) (gaugeNum 5) or SU(2) (gaugeNum 7)
    gauge field (Gradients/Source/TwoHDM.cpp:106-141, 299-319, 401-424;
    WilsonLoops/Source/TwoHDM.cpp:352-386, 496-552, 600-634, 696-754): fields,
    energy densities and all gaugeNum Gauss violations bitwise."""
    tag = f"2hdm_g{gauge_num}" + ("_gen" if gen else "")
    g = np.load(os.path.join(GOLD, f"harness_{tag}.npz"))
    s, v = g["scalar0"].copy(), g["vec0"].copy()
    d = make_desc(model=GFT_MODEL_2HDM, n=(4, 5, 6), pot2=TWO_HDM_POTENTIAL, gen_rep=int(gen),
                  gauge_num=gauge_num, gpp=0.5)
    assert g["viol_0"].shape[1] == gauge_num
    for k in range(6):
        if k % 3 == 0:
            assert np.array_equal(O.energy_density_f32(d, s, v, k), g[f"dens_{k}"]), f"energy densities, step {k}"
        O.evolve(d, s, v, 1, time_iter=k, energy_freq=3, gauss_freq=3)
        if k % 3 == 0:
            assert np.array_equal(O.gauss_site_f32(d, s, v, k), g[f"viol_{k}"]), f"Gauss violations, step {k}"
        if f"scalar_{k}" in g:
            assert np.array_equal(s, g[f"scalar_{k}"]), f"scalar fields differ after step {k}"
            assert np.array_equal(v, g[f"vec_{k}"]), f"link fields differ after step {k}"


@pytest.mark.parametrize("gauge_num,gen", FAMILY, ids=FAMILY_IDS)
def test_oracle_family_gauge_reproduces_reference_program_output(gauge_num, gen):
    """The unmodified reference program with gaugeNum 5 / 7: printed energies and
    all 2 x gaugeNum Gauss columns digit for digit."""
    tag = f"2hdm_g{gauge_num}" + ("_gen" if gen else "")
    g = np.load(os.path.join(GOLD, f"program_{tag}.npz"))
    n = (6, 5, 7)
    N = int(np.prod(n))
    s = O.random_uniform(2 * N * 8, 7, -0.8, 0.8).reshape(2, N, 8)
    v = identity_links_family(N, gauge_num, gen)
    d = make_desc(model=GFT_MODEL_2HDM, n=n, pot2=TWO_HDM_POTENTIAL, gen_rep=int(gen), gauge_num=gauge_num, gpp=0.5)
    ref, obs, _ = O.evolve(d, s, v, 12, energy_freq=3, gauss_freq=2)
    fmt = np.vectorize(lambda x: float("%.6g" % x))
    e = np.array([r[:6] for r, o in zip(ref, obs) if o.has_energy])
    gs = np.array([r[6:] for r, o in zip(ref, obs) if o.has_gauss])
    assert gs.shape[1] == 2 * gauge_num
    assert np.array_equal(fmt(e), g["energy"])
    assert np.array_equal(fmt(gs), g["gauss"])


GLOBAL = [(3, None), (8, TWO_HDM_POTENTIAL)]


def global_desc(n, ns, pot2, d=(0.5, 0.5, 0.5), form=0):
    from gftevolver_amd._abi import GFT_MODEL_GLOBAL
    return make_desc(model=GFT_MODEL_GLOBAL, form=form, n=n, d=d, pot2=pot2, num_scalar=ns,
                     potential=2 if pot2 else 1)


@pytest.mark.parametrize("ns,pot2", GLOBAL, ids=["n3-son", "n8-double-son"])
def test_oracle_global_model_bitwise_vs_reference_harness(ns, pot2):
    """The global O(N) model (Gradients/Source/Global.cpp with the null Wilson
    loop; SO_N with 3 components, Double_SO_N with 8) on unequal spacings:
    fields after steps 1, 3, 6 and the energy densities bitwise."""
    g = np.load(os.path.join(GOLD, f"harness_global_n{ns}.npz"))
    s = g["scalar0"].copy()
    d = global_desc((4, 5, 6), ns, pot2, d=(0.5, 0.4, 0.6))
    for k in range(6):
        if k % 3 == 0:
            assert np.array_equal(O.energy_density_f32(d, s, np.zeros(1, np.float32), k), g[f"dens_{k}"]), k
        O.evolve(d, s, None, 1, time_iter=k)
        if f"scalar_{k}" in g:
            assert np.array_equal(s, g[f"scalar_{k}"]), f"scalar fields differ after step {k}"


@pytest.mark.parametrize("ns,pot2", GLOBAL, ids=["n3-son", "n8-double-son"])
def test_oracle_global_model_reproduces_reference_program_output(ns, pot2):
    g = np.load(os.path.join(GOLD, f"program_global_n{ns}.npz"))
    n = (6, 5, 7)
    N = int(np.prod(n))
    s = O.random_uniform(2 * N * ns, 7, -0.8, 0.8).reshape(2, N, ns)
    ref, obs, _ = O.evolve(global_desc(n, ns, pot2), s, None, 12, energy_freq=3)
    fmt = np.vectorize(lambda x: float("%.6g" % x))
    assert ref.shape[1] == 6
    assert np.array_equal(fmt(ref), g["energy"])


def test_fp64_global_model_free_field_dispersion_and_energy():
    """Closed forms: with lambda = 0 the global model is N free fields; a plane
    wave cos(k.x - w t) with the leapfrog dispersion
    4 sin^2(w dt/2)/dt^2 = sum_a 4 sin^2(k_a dx_a/2)/dx_a^2 stays a plane wave, and
    the FP64 energy is conserved to the leapfrog's O(dt^2) oscillation."""
    n = (8, 6, 10)
    dxs = (0.5, 0.4, 0.6)
    dt = 0.05
    d = global_desc(n, 2, None, d=dxs)
    d.lambda_ = 0.0
    d.dt = dt
    k = [2 * np.pi * m / (ni * dx) for m, ni, dx in zip((1, 2, 1), n, dxs)]
    om = 2 / dt * np.arcsin(dt / 2 * np.sqrt(sum(4 / dx ** 2 * np.sin(ki * dx / 2) ** 2 for ki, dx in zip(k, dxs))))
    x = np.stack(np.meshgrid(*[np.arange(ni) * dx for ni, dx in zip(n, dxs)], indexing="ij"), -1).reshape(-1, 3)
    ph = x @ np.array(k)
    s = np.zeros((2, x.shape[0], 2))
    s[1, :, 0] = np.cos(ph)              # time 0 (current slot at time_iter 0)
    s[0, :, 0] = np.cos(ph + om * dt)    # time -dt
    s[1, :, 1] = 0.5 * np.sin(ph)
    s[0, :, 1] = 0.5 * np.sin(ph + om * dt)
    _, obs, _ = O.evolve(d, s, None, 20, energy_freq=5)
    cur = (20 + 1) % 2
    np.testing.assert_allclose(s[cur, :, 0], np.cos(ph - om * 20 * dt), atol=1e-10)
    np.testing.assert_allclose(s[cur, :, 1], 0.5 * np.sin(ph - om * 20 * dt), atol=1e-10)
    e = np.array([o.energy[0] for o in obs])
    assert e.min() > 0 and np.ptp(e) / e.mean() < 1e-2


@pytest.mark.parametrize("case,mask", [("xz", 1 | 2 | 16 | 32), ("y", 4 | 8)], ids=["xz", "y"])
def test_oracle_reproduces_reference_program_with_fixed_faces(case, mask):
    """Fixed faces (Fixed.cpp: frozen one-site shells, left out of the
    diagnostics): the reference program's printed energies and Gauss
    violations, digit for digit, for x and z fixed and for y fixed."""
    g = np.load(os.path.join(GOLD, "program_sm_fixed.npz"))
    n = (6, 5, 7)
    N = int(np.prod(n))
    s = O.random_uniform(2 * N * 4, 7, -0.8, 0.8).reshape(2, N, 4)
    v = O.identity_links(N)
    d = make_desc(n=n, fixed_faces=mask)
    s0 = s.copy()
    ref, obs, _ = O.evolve(d, s, v, 12, energy_freq=3, gauss_freq=2)
    fmt = np.vectorize(lambda x: float("%.6g" % x))
    e = np.array([r[:6] for r, o in zip(ref, obs) if o.has_energy])
    gs = np.array([r[6:] for r, o in zip(ref, obs) if o.has_gauss])
    assert np.array_equal(fmt(e), g[f"energy_{case}"])
    assert np.array_equal(fmt(gs), g[f"gauss_{case}"])
    # The frozen shell keeps its initial values in both slots.
    frozen = O.frozen_mask(n, mask)
    assert frozen.any() and not frozen.all()
    assert np.array_equal(s[:, frozen], s0[:, frozen])
    # Fully periodic differs: the boundary condition is exercised.
    s2 = O.random_uniform(2 * N * 4, 7, -0.8, 0.8).reshape(2, N, 4)
    ref2, obs2, _ = O.evolve(make_desc(n=n), s2, O.identity_links(N), 12, energy_freq=3, gauss_freq=2)
    assert not np.array_equal(fmt(np.array([r[:6] for r, o in zip(ref2, obs2) if o.has_energy])), g[f"energy_{case}"])


def test_reference_site_order_is_a_permutation_with_shell_first():
    n = (5, 4, 6)
    order = O.site_order(*n)
    assert sorted(order.tolist()) == list(range(np.prod(n)))
    # First region is the -x face: x=0, y in [1,ny-1), z in [1,nz-1) (Lattice.cpp:262-290).
    first = [(0 * 4 + y) * 6 + z for y in range(1, 3) for z in range(1, 5)]
    assert order[:len(first)].tolist() == first
    interior = [(x * 4 + y) * 6 + z for x in range(1, 4) for y in range(1, 3) for z in range(1, 5)]
    assert order[-len(interior):].tolist() == interior


@pytest.mark.parametrize("model,pot2", [(GFT_MODEL_SM, None), (GFT_MODEL_2HDM, TWO_HDM_POTENTIAL)],
                         ids=["sm", "2hdm"])
def test_fp64_oracle_gauge_covariance(model, pot2):
    """evolve(G state) == G evolve(state); energy and |G| gauge invariant."""
    n = (4, 4, 5)
    Ns = 4 if model == GFT_MODEL_SM else 8
    d = make_desc(model=model, n=n, pot2=pot2)
    s, v = L.near_vacuum_state(n, Ns, seed=3, amp=0.2, mom=0.02, link_eps=0.2)
    rng = np.random.default_rng(11)
    N = int(np.prod(n))
    a = rng.uniform(-np.pi, np.pi, N)
    w = L.random_su2(rng, (N,), 3.0)
    s1, v1 = s.copy(), v.copy()
    s2, v2 = L.gauge_transform(s, v, n, a, w)
    _, o1, _ = O.evolve(d, s1, v1, 3, energy_freq=1, gauss_freq=1)
    _, o2, _ = O.evolve(d, s2, v2, 3, energy_freq=1, gauss_freq=1)
    s3, v3 = L.gauge_transform(s1, v1, n, a, w)
    np.testing.assert_allclose(s2, s3, atol=1e-12)
    # Phases are compared modulo 2 pi, quaternions directly.
    dv = (v2 - v3).reshape(2, N, 3, 5)
    dv[..., 0] = (dv[..., 0] + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(dv).max() < 1e-12
    for x, y in zip(o1, o2):
        np.testing.assert_allclose(np.array(x.energy[:]), np.array(y.energy[:]), rtol=1e-11)
        # G^Y is gauge invariant; G^a (a=1..3) is an adjoint vector, so only its
        # per-component sums change under an SU(2) gauge transformation.
        np.testing.assert_allclose(x.gauss_int[0], y.gauss_int[0], rtol=1e-9)


def test_fp64_oracle_time_reversible():
    n = (4, 5, 4)
    d = make_desc(n=n)
    s, v = L.near_vacuum_state(n, 4, seed=5, amp=0.2, mom=0.02, link_eps=0.1)
    s0, v0 = s.copy(), v.copy()
    O.evolve(d, s, v, 10)
    # After 10 steps the current slot is (10+1)%2 = 1; swap slots and run 10 more.
    s, v = s[::-1].copy(), v[::-1].copy()
    O.evolve(d, s, v, 10)
    s, v = s[::-1].copy(), v[::-1].copy()
    np.testing.assert_allclose(s, s0, atol=1e-11)
    np.testing.assert_allclose(v, v0, atol=1e-11)


def test_fp64_oracle_free_field_dispersion():
    """g = g' = lambda = 0: exact plane-wave solution of the leapfrog
    phi(t) = A cos(k.x - w t) with sin(w dt/2) = (dt/2) w_lat,
    w_lat^2 = sum_i 4 sin^2(k_i dx_i / 2) / dx_i^2."""
    n, dxs, dt = (8, 6, 10), (0.5, 0.4, 0.3), 0.05
    d = make_desc(n=n, d=dxs, dt=dt, g=0.0, gp=0.0, lam=0.0)
    m = np.array([1, 2, 3])
    k = 2 * np.pi * m / (np.array(n) * np.array(dxs))
    wl = np.sqrt(sum(4 * np.sin(k[i] * dxs[i] / 2) ** 2 / dxs[i] ** 2 for i in range(3)))
    w = 2 / dt * np.arcsin(dt / 2 * wl)
    X = np.stack(np.meshgrid(*[np.arange(n[i]) * dxs[i] for i in range(3)], indexing="ij"), -1).reshape(-1, 3)
    N = X.shape[0]
    amp = np.array([0.3, -0.1, 0.2, 0.05])

    def phi(t):
        return np.cos(X @ k - w * t)[:, None] * amp[None, :]

    s = np.stack([phi(-dt), phi(0.0)])
    v = O.identity_links(N, np.float64)
    nsteps = 40
    O.evolve(d, s, v, nsteps)
    cur = (nsteps + 1) % 2
    np.testing.assert_allclose(s[cur], phi(nsteps * dt), atol=1e-12)


def _forces(d, s, v):
    """One step from a state with zero momentum: returns (F_phi, F_link[N,3,4])
    from phi(t+dt) = phi + dt^2 F and R = dt^2 F_link."""
    s1 = np.stack([s, s]).copy()
    v1 = np.stack([v, v]).copy()
    O.evolve(d, s1, v1, 1)
    dt2 = d.dt ** 2
    Fphi = (s1[0] - s) / dt2
    vv = v.reshape(-1, 3, 5)
    vn = v1[0].reshape(-1, 3, 5)
    F = np.zeros((vv.shape[0], 3, 4))
    F[..., 0] = np.sin(vn[..., 0] - vv[..., 0]) / dt2
    F[..., 1:] = L.qmul(vn[..., 1:], L.qconj(vv[..., 1:]))[..., 1:] / dt2
    return Fphi, F


def _potential_energy(d, s, v):
    """Potential + gradient + magnetic energy (x dx dy dz) from the FP64 diagnostics."""
    s2 = np.stack([s, s]).copy()
    v2 = np.stack([v, v]).copy()
    _, obs, _ = O.evolve(d, s2, v2, 1, energy_freq=1)
    e = obs[0].energy
    return e[1] + e[2] + e[4]


@pytest.mark.parametrize("model,pot2", [(GFT_MODEL_SM, None), (GFT_MODEL_2HDM, TWO_HDM_POTENTIAL)],
                         ids=["sm", "2hdm"])
def test_fp64_oracle_force_is_minus_energy_gradient(model, pot2):
    """Scalar force = -(1/dV) dH/dphi; U(1) force = -(g'^2 dx^2 / 4 dV) dH/dy;
    SU(2) force along left multiplication by exp(i e sigma^a) = -(g^2 dx^2 / 4 dV) dH/de.
    The masses 4/(g^2 dx^2) follow from the electric energy of the same code."""
    n = (3, 4, 3)
    Ns = 4 if model == GFT_MODEL_SM else 8
    dxs = (0.5, 0.6, 0.45)
    d = make_desc(model=model, n=n, d=dxs, pot2=pot2)
    s, v = L.near_vacuum_state(n, Ns, seed=8, amp=0.25, mom=0.0, link_eps=0.3)
    s, v = s[1], v[1]
    dV = np.prod(dxs)
    Fphi, F = _forces(d, s, v)
    eps = 1e-5
    rng = np.random.default_rng(0)
    for x in rng.choice(s.shape[0], 3, replace=False):
        for c in range(Ns):
            sp, sm = s.copy(), s.copy()
            sp[x, c] += eps
            sm[x, c] -= eps
            dH = (_potential_energy(d, sp, v) - _potential_energy(d, sm, v)) / (2 * eps)
            assert abs(Fphi[x, c] + dH / dV) < 1e-5 * (1 + abs(Fphi[x, c]))
        for i in range(3):
            vp, vm = v.copy().reshape(-1, 3, 5), v.copy().reshape(-1, 3, 5)
            vp[x, i, 0] += eps
            vm[x, i, 0] -= eps
            dH = (_potential_energy(d, s, vp.reshape(-1, 15)) - _potential_energy(d, s, vm.reshape(-1, 15))) / (2 * eps)
            expect = -d.gp ** 2 * dxs[i] ** 2 / (4 * dV) * dH
            assert abs(F[x, i, 0] - expect) < 1e-5 * (1 + abs(expect))
            for a in range(1, 4):
                e = np.zeros(4)
                e[0], e[a] = np.cos(eps), np.sin(eps)
                vp, vm = v.copy().reshape(-1, 3, 5), v.copy().reshape(-1, 3, 5)
                vp[x, i, 1:] = L.qmul(e, vp[x, i, 1:])
                vm[x, i, 1:] = L.qmul(L.qconj(e), vm[x, i, 1:])
                dH = (_potential_energy(d, s, vp.reshape(-1, 15))
                      - _potential_energy(d, s, vm.reshape(-1, 15))) / (2 * eps)
                expect = -d.g ** 2 * dxs[i] ** 2 / (4 * dV) * dH
                assert abs(F[x, i, a] - expect) < 1e-5 * (1 + abs(expect))


def test_fp64_oracle_gauss_law_conserved():
    """From E = 0, pi = 0 (both slots equal) the Gauss law holds exactly and
    the FP64 evolution keeps it at rounding level."""
    n = (6, 5, 4)
    d = make_desc(n=n)
    s, v = L.near_vacuum_state(n, 4, seed=4, amp=0.3, mom=0.0, link_eps=0.3)
    s[0], v[0] = s[1], v[1]
    _, obs, _ = O.evolve(d, s, v, 30, energy_freq=10, gauss_freq=1)
    gi = np.array([o.gauss_int[:o.num_eq] for o in obs if o.has_gauss])
    sc = np.array([o.gauss_scale[:o.num_eq] for o in obs if o.has_gauss])
    assert (gi[5:] / sc[5:]).max() < 1e-12
