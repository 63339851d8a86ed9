"""CPU checks of the drop-in boundary: the HIP library loads (no GPU needed) and
exports every symbol include/gft.h declares; the ctypes structures match the
header's layout."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from gftevolver_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gft.h")
LIB = os.path.join(ROOT, "gftevolver_amd", "libgft.so")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gft_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("gft_create", "gft_destroy", "gft_evolve", "gft_import_two_slot_f32",
                     "gft_export_two_slot_f64", "gft_last_error", "gft_last_step_ms"):
        assert required in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgft.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # gft_obs: 4 unsigned then 6 + 3 * GFT_MAX_EQ doubles.
    max_eq = int(re.search(r"#define GFT_MAX_EQ (\d+)", open(HEADER).read()).group(1))
    assert max_eq == _abi.MAX_EQ
    assert C.sizeof(_abi.GftObs) == 16 + (6 + 3 * max_eq) * 8
    assert C.sizeof(_abi.GftLoops) == 5 * 8
    fields = [f[0] for f in _abi.GftDesc._fields_]
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = txt[txt.index("typedef struct gft_desc"):txt.index("} gft_desc;")]
    hdr = re.findall(r"\b(?:int|double|float|unsigned)\s+([^;]+);", body)
    names = [x.strip() for grp in hdr for x in grp.split(",")]
    assert [n if n != "lambda" else "lambda_" for n in names] == fields


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgft.so not built")
@pytest.mark.parametrize("faces,msg", [(1, "both faces of an axis (axis 0)"), (4 | 16 | 32, "(axis 1)"),
                                       (64, "six face bits"), (-1, "six face bits")])
def test_create_rejects_invalid_face_masks_before_touching_a_device(faces, msg):
    """gft_create validates the configuration before any HIP call, with the
    reference's message for an unpaired periodic face (Lattice.cpp:244-259)."""
    lib = C.CDLL(LIB)
    lib.gft_last_error.restype = C.c_char_p
    lib.gft_last_error.argtypes = [C.c_void_p]
    d = _abi.make_desc(form=_abi.GFT_FORM_MOMENTUM_F32, n=(8, 8, 8), fixed_faces=faces)
    ctx = C.c_void_p()
    assert lib.gft_create(C.byref(d), C.byref(ctx)) == -1
    assert not ctx.value
    assert msg in lib.gft_last_error(None).decode()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgft.so not built")
@pytest.mark.parametrize("model,form,gauge_num,msg", [
    (_abi.GFT_MODEL_SM, _abi.GFT_FORM_MOMENTUM_F32, 5, "two doublets"),
    (_abi.GFT_MODEL_SM, _abi.GFT_FORM_TWO_SLOT_F64, 7, "two doublets"),
    (_abi.GFT_MODEL_2HDM, _abi.GFT_FORM_TWO_SLOT_F64, 6, "gauge_num is 4")])
def test_create_rejects_unsupported_gauge_groups_before_touching_a_device(model, form, gauge_num, msg):
    lib = C.CDLL(LIB)
    lib.gft_last_error.restype = C.c_char_p
    lib.gft_last_error.argtypes = [C.c_void_p]
    d = _abi.make_desc(model=model, form=form, n=(8, 8, 8), pot2=_abi.TWO_HDM_POTENTIAL, gauge_num=gauge_num)
    ctx = C.c_void_p()
    assert lib.gft_create(C.byref(d), C.byref(ctx)) == -1
    assert msg in lib.gft_last_error(None).decode()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgft.so not built")
@pytest.mark.parametrize("form,ns,potential,msg", [
    (_abi.GFT_FORM_MOMENTUM_F32, 4, 1, "two-slot forms"),
    (_abi.GFT_FORM_TWO_SLOT_F64, 9, 1, "1 to 8 scalar components"),
    (_abi.GFT_FORM_TWO_SLOT_F64, 4, 2, "Double_SO_N, 8 components")])
def test_create_rejects_unsupported_global_models_before_touching_a_device(form, ns, potential, msg):
    lib = C.CDLL(LIB)
    lib.gft_last_error.restype = C.c_char_p
    lib.gft_last_error.argtypes = [C.c_void_p]
    d = _abi.make_desc(model=_abi.GFT_MODEL_GLOBAL, form=form, n=(8, 8, 8), num_scalar=ns, potential=potential)
    ctx = C.c_void_p()
    assert lib.gft_create(C.byref(d), C.byref(ctx)) == -1
    assert msg in lib.gft_last_error(None).decode()
