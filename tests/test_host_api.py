"""The C++ host API (gftevolver_amd/host) as a drop-in for the reference program:
it reads the reference's own positional .cfg files and writes the reference's
Data/global_quantities*.dat.

CPU: the config reader against values written into the shipped config files.
GPU: gft_run on the configuration of the golden program run reproduces the
reference program's energy file (the reference accumulates in float, the GPU
in FP64, so 2e-5 relative) and a Gauss violation no larger than the
reference's own float-rounding level."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from gftevolver_amd._abi import TWO_HDM_POTENTIAL
from tests import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GFT_RUN = os.path.join(ROOT, "gftevolver_amd", "gft_run")
TEMPLATE = os.path.join(ROOT, "oracle", "_ref", "cfg")
sys.path.insert(0, os.path.join(ROOT, "oracle"))

needs_cfg = pytest.mark.skipif(not (os.path.isdir(TEMPLATE) and os.path.exists(GFT_RUN)),
                               reason="reference config template or gft_run not built")


def _prepare(td, model, n, nt, ns, efreq=3, gfreq=2, faces=(3, 3, 3, 3, 3, 3), gen=False, nvd=None):
    import refcfg
    refcfg.prepare(td)
    refcfg.write_model(td, model, 0.65, 0.35, pot2=TWO_HDM_POTENTIAL if model == 1 else None, gen=gen)
    refcfg.write_lattice(td, n, (0.5, 0.5, 0.5), nt, 0.05, ns, ic_seed=7, ic_range=(-0.8, 0.8), efreq=efreq,
                         gfreq=gfreq, faces=faces, nvd=nvd or (4 if gen else 5))
    return os.path.join(td, "cfg")


@needs_cfg
@pytest.mark.parametrize("model", [0, 1], ids=["sm", "2hdm"])
def test_config_reader_matches_values_written(tmp_path, model):
    ns = 4 if model == 0 else 8
    root = _prepare(str(tmp_path), model, (6, 5, 7), 12, ns)
    out = subprocess.run([GFT_RUN, root, str(tmp_path / "data"), "dump_config"], capture_output=True, text=True,
                         check=True).stdout
    kv = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
    assert kv["n"] == ["6", "5", "7"] and float(kv["dt"][0]) == 0.05 and kv["nt"] == ["12"]
    assert kv["ns"] == [str(ns)] and kv["nv"] == ["15"] and kv["model"] == [str(model)]
    assert float(kv["g"][0]) == 0.65 and float(kv["gp"][0]) == 0.35
    # The default form is the fast path (GFT_FORM_MOMENTUM_F32 = 2) for both models.
    assert kv["form"] == ["2"]
    if model == 0:
        assert float(kv["lambda"][0]) == 0.13 and float(kv["eta_sqr"][0]) == 1.0
    else:
        p = TWO_HDM_POTENTIAL
        assert float(kv["m1"][0]) == p["m1"] and float(kv["l3"][0]) == p["l3"]
        assert float(kv["m12r"][0]) == p["m12r"] and float(kv["l4m5"][0]) == p["l4m5"]
    # RandomUniform initial state: same generator and sequence as the oracle's restatement.
    N = 6 * 5 * 7
    ref = O.random_uniform(2 * N * ns, 7, -0.8, 0.8)
    assert abs(float(kv["scalar_sum"][0]) - float(ref.astype(np.float64).sum())) < 1e-9


@needs_cfg
def test_config_reader_fixed_faces(tmp_path):
    """Face types 1 (fixed) and 3 (periodic) map to gft_desc::fixed_faces bits;
    a periodic face must be paired with a periodic opposite face."""
    root = _prepare(str(tmp_path / "a"), 0, (6, 5, 7), 12, 4, faces=(1, 1, 3, 3, 1, 1))
    out = subprocess.run([GFT_RUN, root, str(tmp_path / "a" / "data"), "dump_config"], capture_output=True, text=True,
                         check=True).stdout
    kv = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
    assert kv["fixed_faces"] == [str(1 | 2 | 16 | 32)]
    root = _prepare(str(tmp_path / "b"), 0, (6, 5, 7), 12, 4, faces=(1, 3, 3, 3, 3, 3))
    r = subprocess.run([GFT_RUN, root, str(tmp_path / "b" / "data"), "dump_config"], capture_output=True, text=True)
    assert r.returncode != 0 and "must be assigned to both faces of an axis" in r.stderr


@needs_cfg
@pytest.mark.parametrize("n,faces", [((6, 5, 7), (3, 3, 3, 3, 3, 3)), ((6, 5, 7), (1, 1, 3, 3, 1, 1)),
                                     ((2, 3, 5), (1, 1, 1, 1, 1, 1)), ((4, 2, 3), (3, 3, 1, 1, 3, 3))],
                         ids=["periodic", "fixed-xz", "thin-fixed", "thin-mixed"])
def test_boundary_regions_match_reference_order_and_frozen_shell(tmp_path, n, faces):
    """Lattice::boundaryConditions: the non-empty regions, listed in order and
    followed by the interior, visit every site once in the order of the oracle's
    restatement of the reference loops (Lattice.cpp:262-355, 706-716,
    BoundaryCondition.cpp:21-70); the Fixed regions hold exactly the sites the
    GPU kernels keep frozen (a site on the shell of any fixed face)."""
    root = _prepare(str(tmp_path), 0, n, 12, 4, faces=faces)
    out = subprocess.run([GFT_RUN, root, str(tmp_path / "data"), "dump_config"], capture_output=True, text=True,
                         check=True).stdout
    nx, ny, nz = n
    order, fixed = [], set()
    lines = [l.split() for l in out.splitlines()]
    regions = [[int(v) for v in l[1:]] for l in lines if l[0] == "region"]
    inner = [[int(v) for v in l[1:]] for l in lines if l[0] == "interior"][0]

    def sites(lim):
        return [(x * ny + y) * nz + z for x in range(lim[0], lim[1]) for y in range(lim[2], lim[3])
                for z in range(lim[4], lim[5])]

    for r in regions:
        assert r[3] in (1, 3)
        s = sites(r[4:])
        assert s, "empty regions are dropped"
        order += s
        if r[3] == 1:
            fixed.update(s)
            assert any(r[a] != 0 and faces[2 * a + (r[a] > 0)] == 1 for a in range(3))
    order += sites(inner)
    assert np.array_equal(np.array(order), O.site_order(*n))
    idx = np.arange(nx * ny * nz)
    x, y, z = idx // (ny * nz), (idx // nz) % ny, idx % nz
    frozen = np.zeros(idx.size, bool)
    for f, (c, v) in enumerate(((x, 0), (x, nx - 1), (y, 0), (y, ny - 1), (z, 0), (z, nz - 1))):
        if faces[f] == 1:
            frozen |= c == v
    assert fixed == set(idx[frozen].tolist())
    kv = {l[0]: l[1:] for l in lines}
    assert int(kv["fixed_faces"][0]) == sum(1 << f for f in range(6) if faces[f] == 1)


@needs_cfg
@pytest.mark.parametrize("model", [0, 1], ids=["sm", "2hdm"])
def test_config_reader_generator_form(tmp_path, model):
    """4 vector components per direction select the generator form of the SU(2)
    links (SM: by the count, StandardModel.cpp:206-209; 2HDM: TwoHDM.cfg's flag
    with the count fixing gaugeNum, TwoHDM.cpp:215-242)."""
    ns = 4 if model == 0 else 8
    root = _prepare(str(tmp_path / "a"), model, (6, 5, 7), 12, ns, gen=True)
    out = subprocess.run([GFT_RUN, root, str(tmp_path / "a" / "data"), "dump_config"], capture_output=True, text=True,
                         check=True).stdout
    kv = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
    assert kv["nv"] == ["12"] and kv["gen_rep"] == ["1"]
    # 5 components per direction with the 2HDM generator flag set: gaugeNum 5.
    if model == 1:
        root = _prepare(str(tmp_path / "b"), model, (6, 5, 7), 12, ns, gen=True, nvd=5)
        out = subprocess.run([GFT_RUN, root, str(tmp_path / "b" / "data"), "dump_config"], capture_output=True,
                             text=True, check=True).stdout
        kv = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
        assert kv["gauge_num"] == ["5"] and kv["gen_rep"] == ["1"] and float(kv["gpp"][0]) == 0.5
    # A count neither form allows: the reference's message.
    root = _prepare(str(tmp_path / "c"), model, (6, 5, 7), 12, ns, gen=model == 1, nvd=3)
    r = subprocess.run([GFT_RUN, root, str(tmp_path / "c" / "data"), "dump_config"], capture_output=True, text=True)
    assert r.returncode != 0 and "9 have been assigned" in r.stderr


@needs_cfg
def test_config_errors_keep_reference_messages(tmp_path):
    import refcfg
    root = _prepare(str(tmp_path), 0, (6, 5, 7), 12, 4)
    refcfg.set_values(str(tmp_path), "Core", "Model.cfg", [7, 2, 1, 0.0, 0, 0, 0.0])
    r = subprocess.run([GFT_RUN, root, str(tmp_path / "data"), "dump_config"], capture_output=True, text=True)
    assert r.returncode != 0 and "The chosen potential type (7) is invalid." in r.stderr


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("model,tag", [(0, "sm"), (1, "2hdm")], ids=["sm", "2hdm"])
def test_gft_run_reproduces_reference_program_output(tmp_path, model, tag):
    g = np.load(os.path.join(ROOT, "tests", "golden", f"program_{tag}.npz"))
    ns = 4 if model == 0 else 8
    root = _prepare(str(tmp_path), model, (6, 5, 7), 12, ns)
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, root, str(data), "two_slot_f32"], check=True, capture_output=True, timeout=120)
    E = np.loadtxt(data / "global_quantities.dat", ndmin=2)
    G = np.loadtxt(data / "global_quantities_gauge.dat", ndmin=2)
    assert E.shape == g["energy"].shape and G.shape == g["gauss"].shape
    np.testing.assert_allclose(E, g["energy"], rtol=2e-5, atol=1e-5 * np.abs(g["energy"]).max())
    assert (G <= 10 * g["gauss"] + 1e-7).all()


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("model,tag", [(0, "sm"), (1, "2hdm")], ids=["sm", "2hdm"])
def test_gft_run_default_fast_path_reproduces_reference_program(tmp_path, model, tag):
    """gft_run with no form argument runs the FP32 momentum form (the fast path)
    for the SM and the 2HDM: it reorders the reference's arithmetic, so the
    energies agree to the same 2e-5 as the float reference's own accumulation
    error and the Gauss violation stays at the reference's level."""
    g = np.load(os.path.join(ROOT, "tests", "golden", f"program_{tag}.npz"))
    ns = 4 if model == 0 else 8
    root = _prepare(str(tmp_path), model, (6, 5, 7), 12, ns)
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, root, str(data)], check=True, capture_output=True, timeout=120)
    E = np.loadtxt(data / "global_quantities.dat", ndmin=2)
    G = np.loadtxt(data / "global_quantities_gauge.dat", ndmin=2)
    assert E.shape == g["energy"].shape and G.shape == g["gauss"].shape
    err = np.abs(E - g["energy"]).max() / np.abs(g["energy"]).max()
    print(f"{tag}: default form, energies vs the reference program: max rel err {err:.2e}")
    np.testing.assert_allclose(E, g["energy"], rtol=2e-5, atol=1e-5 * np.abs(g["energy"]).max())
    assert (G <= 10 * g["gauss"] + 1e-7).all()


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("case,faces", [("xz", (1, 1, 3, 3, 1, 1)), ("y", (3, 3, 1, 1, 3, 3))], ids=["xz", "y"])
def test_gft_run_reproduces_reference_program_with_fixed_faces(tmp_path, case, faces):
    g = np.load(os.path.join(ROOT, "tests", "golden", "program_sm_fixed.npz"))
    root = _prepare(str(tmp_path), 0, (6, 5, 7), 12, 4, faces=faces)
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, root, str(data), "two_slot_f32"], check=True, capture_output=True, timeout=120)
    E = np.loadtxt(data / "global_quantities.dat", ndmin=2)
    G = np.loadtxt(data / "global_quantities_gauge.dat", ndmin=2)
    ge, gg = g[f"energy_{case}"], g[f"gauss_{case}"]
    assert E.shape == ge.shape and G.shape == gg.shape
    np.testing.assert_allclose(E, ge, rtol=2e-5, atol=1e-5 * np.abs(ge).max())
    # With frozen shells the violation is physical (no current through frozen links), so it is
    # compared, not bounded: the GPU sums in FP64 what the reference sums in float.
    np.testing.assert_allclose(G, gg, rtol=1e-3, atol=1e-7)


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("model,tag", [(0, "sm"), (1, "2hdm")], ids=["sm", "2hdm"])
def test_gft_run_reproduces_reference_program_generator_form(tmp_path, model, tag):
    """The reference program with generator-form links (12 vector components)."""
    g = np.load(os.path.join(ROOT, "tests", "golden", f"program_{tag}_gen.npz"))
    ns = 4 if model == 0 else 8
    root = _prepare(str(tmp_path), model, (6, 5, 7), 12, ns, gen=True)
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, root, str(data), "two_slot_f32"], check=True, capture_output=True, timeout=120)
    E = np.loadtxt(data / "global_quantities.dat", ndmin=2)
    G = np.loadtxt(data / "global_quantities_gauge.dat", ndmin=2)
    assert E.shape == g["energy"].shape and G.shape == g["gauss"].shape
    np.testing.assert_allclose(E, g["energy"], rtol=2e-5, atol=1e-5 * np.abs(g["energy"]).max())
    assert (G <= 10 * g["gauss"] + 1e-7).all()


@needs_cfg
@pytest.mark.parametrize("nvd,gen,gauge_num", [(6, False, 5), (9, False, 7), (5, True, 5), (7, True, 7)],
                         ids=["g5", "g7", "g5-gen", "g7-gen"])
def test_config_reader_family_gauge(tmp_path, nvd, gen, gauge_num):
    """The component count and the generator flag fix the 2HDM's gauge group
    (TwoHDM.cpp:215-242); the Higgs-family coupling comes from TwoHDM.cfg."""
    root = _prepare(str(tmp_path), 1, (6, 5, 7), 12, 8, gen=gen, nvd=nvd)
    out = subprocess.run([GFT_RUN, root, str(tmp_path / "data"), "dump_config"], capture_output=True, text=True,
                         check=True).stdout
    kv = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
    assert kv["gauge_num"] == [str(gauge_num)] and kv["nv"] == [str(3 * nvd)] and kv["gen_rep"] == [str(int(gen))]
    # Family gauge fields run in the fast path (GFT_FORM_MOMENTUM_F32 = 2) by default.
    assert kv["form"] == ["2"]


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("form", ["two_slot_f32", None], ids=["two_slot_f32", "default"])
@pytest.mark.parametrize("gauge_num,gen", [(5, False), (5, True), (7, False), (7, True)],
                         ids=["g5", "g5-gen", "g7", "g7-gen"])
def test_gft_run_reproduces_reference_program_family_gauge(tmp_path, gauge_num, gen, form):
    """The reference program's 2HDM with a Higgs-family gauge field: energies and
    all 2 x gaugeNum Gauss columns, in the reference-order FP32 two-slot form and
    in the default form (the FP32 momentum form, the fast path), whose reordered
    arithmetic agrees to the float reference's own accumulation error."""
    tag = f"2hdm_g{gauge_num}" + ("_gen" if gen else "")
    g = np.load(os.path.join(ROOT, "tests", "golden", f"program_{tag}.npz"))
    nvd = (4 if gen else 5) + (1 if gauge_num == 5 else (3 if gen else 4))
    root = _prepare(str(tmp_path), 1, (6, 5, 7), 12, 8, gen=gen, nvd=nvd)
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, root, str(data)] + ([form] if form else []), check=True, capture_output=True,
                   timeout=120)
    E = np.loadtxt(data / "global_quantities.dat", ndmin=2)
    G = np.loadtxt(data / "global_quantities_gauge.dat", ndmin=2)
    assert E.shape == g["energy"].shape and G.shape == g["gauss"].shape == (G.shape[0], 2 * gauge_num)
    np.testing.assert_allclose(E, g["energy"], rtol=2e-5, atol=1e-5 * np.abs(g["energy"]).max())
    assert (G <= 10 * g["gauss"] + 1e-7).all()


@needs_cfg
@pytest.mark.parametrize("ns,pot2", [(3, None), (8, TWO_HDM_POTENTIAL)], ids=["n3-son", "n8-double-son"])
def test_config_reader_global_model(tmp_path, ns, pot2):
    """Potential 1 or 2, gradient 1 (global), Wilson loop 0 (null): the global
    O(N) model with the Lattice.cfg scalar component count."""
    import refcfg
    refcfg.prepare(str(tmp_path))
    refcfg.write_model(str(tmp_path), 2, 0.65, 0.35, pot2=pot2)
    refcfg.write_lattice(str(tmp_path), (6, 5, 7), (0.5, 0.5, 0.5), 12, 0.05, ns, ic_seed=7, ic_range=(-0.8, 0.8),
                         efreq=3, gfreq=2, nvd=1, gauss=False)
    out = subprocess.run([GFT_RUN, os.path.join(str(tmp_path), "cfg"), str(tmp_path / "data"), "dump_config"],
                         capture_output=True, text=True, check=True).stdout
    kv = {l.split()[0]: l.split()[1:] for l in out.strip().splitlines()}
    assert kv["model"] == ["2"] and kv["ns"] == [str(ns)] and kv["form"] == ["1"]


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("ns,pot2", [(3, None), (8, TWO_HDM_POTENTIAL)], ids=["n3-son", "n8-double-son"])
def test_gft_run_reproduces_reference_program_global_model(tmp_path, ns, pot2):
    import refcfg
    g = np.load(os.path.join(ROOT, "tests", "golden", f"program_global_n{ns}.npz"))
    refcfg.prepare(str(tmp_path))
    refcfg.write_model(str(tmp_path), 2, 0.65, 0.35, pot2=pot2)
    refcfg.write_lattice(str(tmp_path), (6, 5, 7), (0.5, 0.5, 0.5), 12, 0.05, ns, ic_seed=7, ic_range=(-0.8, 0.8),
                         efreq=3, gfreq=2, nvd=1, gauss=False)
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, os.path.join(str(tmp_path), "cfg"), str(data)], check=True, capture_output=True,
                   timeout=120)
    E = np.loadtxt(data / "global_quantities.dat", ndmin=2)
    assert E.shape == g["energy"].shape
    np.testing.assert_allclose(E, g["energy"], rtol=2e-5, atol=1e-5 * np.abs(g["energy"]).max())


def _dump_values(text):
    rows = [np.array(line.split(), dtype=np.float64) for line in str(text).splitlines()]
    assert len({r.size for r in rows}) == 1
    return np.stack(rows)


@needs_cfg
@pytest.mark.gpu
@pytest.mark.parametrize("form", ["two_slot_f32", None], ids=["two_slot_f32", "default"])
@pytest.mark.parametrize("case,both", [("both", 1), ("cur", 0)], ids=["both_slots", "current_slot"])
def test_gft_run_field_dumps_match_reference_program(tmp_path, form, case, both):
    """The OutputFields analyser (OutputFields.cpp:63-133, 157-202) with every
    dump on: the initial fields, continual dumps every 5 completed steps and the
    final fields, both time slots or the current one (slot (t + 1) % 2 after t
    steps).  gft_run writes the same files, the same lines and the same
    components per line as the reference program; the values, printed to 6
    significant digits by both, agree to 1e-5 (the last printed digit plus the
    difference of the FP32 arithmetic orders over 12 steps)."""
    import refcfg
    g = np.load(os.path.join(ROOT, "tests", "golden", "program_sm_dumps.npz"))
    root = _prepare(str(tmp_path), 0, (6, 5, 7), 12, 4)
    refcfg.write_lattice(str(tmp_path), (6, 5, 7), (0.5, 0.5, 0.5), 12, 0.05, 4, ic_seed=7, ic_range=(-0.8, 0.8),
                         efreq=3, gfreq=2, output_fields=True)
    refcfg.set_values(str(tmp_path), "Analysers", "OutputFields.cfg",
                      [both, 1, 1, 1, 5, "initial_fields.dat", "Continual/fields", "final_fields.dat"])
    data = tmp_path / "data"
    subprocess.run([GFT_RUN, root, str(data)] + ([form] if form else []), check=True, capture_output=True,
                   timeout=120)
    names = [str(x) for x in g[f"{case}_names"]]
    written = sorted(os.path.join("Continual", f) for f in os.listdir(data / "Continual"))
    assert written == sorted(n for n in names if n.startswith("Continual"))
    for i, nm in enumerate(names):
        ref = _dump_values(g[f"{case}_{i}"])
        got = _dump_values((data / nm).read_text())
        assert got.shape == ref.shape == (6 * 5 * 7 * (2 if both else 1), 4 + 15), (nm, got.shape, ref.shape)
        err = np.abs(got - ref).max()
        print(f"{form or 'default'} {case} {nm}: max abs diff {err:.2e}")
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5, err_msg=nm)
