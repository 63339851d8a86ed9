"""ctypes mirror of include/gft.h (structures and constants)."""
from __future__ import annotations

import ctypes as C

GFT_MODEL_SM = 0
GFT_MODEL_2HDM = 1
GFT_MODEL_GLOBAL = 2

GFT_FORM_TWO_SLOT_F64 = 0
GFT_FORM_TWO_SLOT_F32 = 1
GFT_FORM_MOMENTUM_F32 = 2
GFT_FORM_MOMENTUM_F64 = 3
GFT_FORM_MOMENTUM_F32_PLAIN = 4

GFT_OK = 0
GFT_EINVAL = -1
GFT_EHIP = -2
GFT_ENOMEM = -3
GFT_EDOMAIN = -4
GFT_ENODEV = -5

NV = 15  # link reals per site: 3 directions x [y, c0, c1, c2, c3]


class GftDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("form", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("g", C.c_double), ("gp", C.c_double), ("div_tol", C.c_double),
        ("lambda_", C.c_double), ("eta_sqr", C.c_double),
        ("m1", C.c_double), ("m2", C.c_double), ("l1", C.c_double), ("l2", C.c_double), ("l3", C.c_double),
        ("break_to_son", C.c_int), ("break_to_uhalfn", C.c_int),
        ("m12r", C.c_double), ("l4p5", C.c_double), ("l6r", C.c_double), ("l7r", C.c_double),
        ("m12i", C.c_double), ("l4m5", C.c_double), ("l5i", C.c_double), ("l6i", C.c_double), ("l7i", C.c_double),
        ("pot_norm", C.c_int), ("pot_offset", C.c_double),
        ("damping", C.c_float), ("nt_damped", C.c_uint),
        ("fixed_faces", C.c_int),
        ("gen_rep", C.c_int),
        ("gauge_num", C.c_int),
        ("gpp", C.c_double),
        ("num_scalar", C.c_int),
        ("potential", C.c_int),
    ]


def link_reals(desc) -> int:
    """Link reals per site of the two-slot arrays: 3 directions x [y, SU(2) (4 as
    a quaternion, 3 as generators), Higgs-family part (gaugeNum 5: y'; 7: a
    second SU(2))]; none for the global O(N) model."""
    if getattr(desc, "model", 0) == GFT_MODEL_GLOBAL:
        return 0
    base = 4 if desc.gen_rep else 5
    gn = desc.gauge_num if desc.gauge_num in (5, 7) else 4
    extra = {4: 0, 5: 1, 7: 3 if desc.gen_rep else 4}[gn]
    return 3 * (base + extra)


MAX_EQ = 7  # GFT_MAX_EQ: Gauss generators of the largest gauge group (gaugeNum 7)


class GftObs(C.Structure):
    _fields_ = [
        ("time_iter", C.c_uint), ("has_energy", C.c_uint), ("has_gauss", C.c_uint), ("num_eq", C.c_uint),
        ("energy", C.c_double * 6),
        ("gauss_int", C.c_double * MAX_EQ), ("gauss_max", C.c_double * MAX_EQ), ("gauss_scale", C.c_double * MAX_EQ),
    ]


class GftLoops(C.Structure):
    _fields_ = [("w11_su2", C.c_double), ("w11_u1", C.c_double), ("w22_su2", C.c_double),
                ("w22_u1", C.c_double), ("phi_sq", C.c_double)]


def make_desc(model=GFT_MODEL_SM, form=GFT_FORM_TWO_SLOT_F64, n=(8, 8, 8), d=(0.5, 0.5, 0.5), dt=0.05,
              g=0.65, gp=0.35, lam=0.13, eta_sqr=1.0, pot2=None, damping=0.0, nt_damped=0, div_tol=1e-9,
              fixed_faces=0, gen_rep=0, gauge_num=4, gpp=0.5, num_scalar=0, potential=0):
    """A gft_desc with the SM couplings of the 32^3 / 256^3 workloads by default.

    pot2: Double_SO_N parameters for the 2HDM (keys as in gft_desc).
    fixed_faces: bit 2a / 2a+1 fixes the low / high face of axis a (x, y, z).
    gen_rep: 1 stores SU(2) links as generator components (12 reals per site).
    gauge_num: 4 (SU(2) x U(1)), 5 or 7 (2HDM with a Higgs-family U(1) or SU(2));
    gpp: the Higgs-family coupling.
    num_scalar, potential: the global O(N) model's N and potential (1 SO_N, 2 Double_SO_N)."""
    ds = GftDesc()
    ds.model, ds.form = model, form
    ds.nx, ds.ny, ds.nz = n
    ds.dx, ds.dy, ds.dz = d
    ds.dt = dt
    ds.g, ds.gp, ds.div_tol = g, gp, div_tol
    ds.lambda_, ds.eta_sqr = lam, eta_sqr
    ds.break_to_son = ds.break_to_uhalfn = 0
    if pot2:
        for k, v in pot2.items():
            setattr(ds, k, v)
    ds.damping, ds.nt_damped = damping, nt_damped
    ds.fixed_faces = fixed_faces
    ds.gen_rep = gen_rep
    ds.gauge_num = gauge_num
    ds.gpp = gpp
    ds.num_scalar = num_scalar
    ds.potential = potential
    return ds


# 2HDM potential of the 192^3 workload in Double_SO_N parameters:
# Phi_k = (f_k0 + i f_k1, f_k2 + i f_k3)/sqrt(2), so Phi1^dag Phi2 = (D + iA)/2 with
# D = f1.f2, A = f1^T J f2; then m1 = -m11^2/2, l1 = lambda1/8, l3 = lambda3/4,
# l4p5 = (lambda4+lambda5)/4, l4m5 = (lambda4-lambda5)/4, m12r = m12^2.
TWO_HDM_POTENTIAL = dict(m1=0.25, m2=0.25, l1=0.0325, l2=0.0325, l3=0.05, break_to_son=1, break_to_uhalfn=1,
                         m12r=0.1, l4p5=-0.0125, l6r=0.0, l7r=0.0, m12i=0.0, l4m5=-0.0375, l5i=0.0,
                         l6i=0.0, l7i=0.0, pot_norm=0, pot_offset=0.0)


class GftMlpDesc(C.Structure):
    _fields_ = [("n_in", C.c_int), ("width", C.c_int), ("n_hidden", C.c_int), ("n_heads", C.c_int),
                ("n_out", C.c_int), ("max_batch", C.c_int)]


class GftGaussDesc(C.Structure):
    _fields_ = [("seed", C.c_ulonglong), ("sigma", C.c_double), ("vev", C.c_double * 2)]


class GftGrfDesc(C.Structure):
    _fields_ = [("seed", C.c_ulonglong), ("rms", C.c_double), ("slope", C.c_double), ("kmax", C.c_double),
                ("vev", C.c_double * 2)]
