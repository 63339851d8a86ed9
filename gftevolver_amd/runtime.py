"""ctypes binding of libgft.so (the C ABI in include/gft.h).

This is the Python caller of the drop-in boundary used by the tests and the
benchmark; the C++ host API (gftevolver_amd/host) is the other caller.  There
is no fallback: if the HIP library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import (GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_FORM_TWO_SLOT_F32, GFT_FORM_TWO_SLOT_F64,
                   GftDesc, GftGaussDesc, GftGrfDesc, GftLoops, GftMlpDesc, GftObs, link_reals)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgft.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")

_L = C.CDLL(LIB_PATH)

_P = C.POINTER
_L.gft_create.argtypes = [_P(GftDesc), _P(C.c_void_p)]
_L.gft_destroy.argtypes = [C.c_void_p]
_L.gft_destroy.restype = None
_L.gft_last_error.argtypes = [C.c_void_p]
_L.gft_last_error.restype = C.c_char_p
_L.gft_device_info.argtypes = [C.c_char_p, C.c_size_t, _P(C.c_int)]
for _n, _t in (("f32", C.c_float), ("f64", C.c_double)):
    getattr(_L, f"gft_import_two_slot_{_n}").argtypes = [C.c_void_p, _P(_t), _P(_t), C.c_uint]
    getattr(_L, f"gft_export_two_slot_{_n}").argtypes = [C.c_void_p, _P(_t), _P(_t), _P(C.c_uint)]
_L.gft_evolve.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_uint, _P(GftObs), C.c_uint, _P(C.c_uint)]
_L.gft_observe.argtypes = [C.c_void_p, _P(GftObs)]
_L.gft_observe_noalloc.argtypes = [C.c_void_p, _P(GftObs)]
_L.gft_wilson_loops.argtypes = [C.c_void_p, _P(GftLoops)]
_L.gft_sync.argtypes = [C.c_void_p]
_L.gft_set_launch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
_L.gft_bench_copy.argtypes = [C.c_size_t, C.c_int, _P(C.c_double)]
_L.gft_bench_streams.argtypes = [C.c_void_p, C.c_int, _P(C.c_double)]
_L.gft_last_step_ms.argtypes = [C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_uint)]
_L.gft_ensemble_create.argtypes = [_P(GftDesc), C.c_int, _P(C.c_void_p)]
_L.gft_ensemble_destroy.argtypes = [C.c_void_p]
_L.gft_ensemble_destroy.restype = None
_L.gft_ensemble_last_error.argtypes = [C.c_void_p]
_L.gft_ensemble_last_error.restype = C.c_char_p
_L.gft_ensemble_member.argtypes = [C.c_void_p, C.c_int, _P(C.c_void_p)]
_L.gft_ensemble_set_launch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
_L.gft_ensemble_evolve.argtypes = [C.c_void_p, C.c_uint, C.c_int]
_L.gft_ensemble_last_step_ms.argtypes = [C.c_void_p, _P(C.c_double)]
_L.gft_init_grf.argtypes = [C.c_void_p, _P(GftGrfDesc)]
_L.gft_init_gaussian.argtypes = [C.c_void_p, _P(GftGaussDesc)]
_L.gft_checkpoint_size.argtypes = [C.c_void_p, _P(C.c_size_t)]
_L.gft_checkpoint.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
_L.gft_restore.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
_L.gft_snapshot.argtypes = [C.c_void_p]
_L.gft_rollback.argtypes = [C.c_void_p, _P(C.c_uint)]
_F = _P(C.c_float)
_L.gft_mlp_create.argtypes = [_P(GftMlpDesc), _P(C.c_void_p)]
_L.gft_mlp_destroy.argtypes = [C.c_void_p]
_L.gft_mlp_destroy.restype = None
_L.gft_mlp_last_error.argtypes = [C.c_void_p]
_L.gft_mlp_last_error.restype = C.c_char_p
_L.gft_mlp_num_params.argtypes = [C.c_void_p]
_L.gft_mlp_num_params.restype = C.c_longlong
_L.gft_mlp_set_params.argtypes = [C.c_void_p, _F]
_L.gft_mlp_get_params.argtypes = [C.c_void_p, _F]
_L.gft_mlp_forward.argtypes = [C.c_void_p, _F, C.c_int, _F]
_L.gft_mlp_gradients.argtypes = [C.c_void_p, _F, _F, _F, C.c_int, _F, _F]
_L.gft_mlp_train.argtypes = [C.c_void_p, _F, _F, _F, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, _F]


class GftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gft error {code}: {msg}")
        self.code = code


def lib():
    return _L


def device_info():
    buf = C.create_string_buffer(128)
    cu = C.c_int(0)
    rc = _L.gft_device_info(buf, 128, C.byref(cu))
    if rc != 0:
        raise GftError(rc, "no HIP device")
    return buf.value.decode(), cu.value


def bench_copy(nbytes: int, reps: int = 5) -> float:
    """Average ms of a plain device copy of nbytes (read nbytes, write nbytes)."""
    ms = C.c_double()
    rc = _L.gft_bench_copy(nbytes, reps, C.byref(ms))
    if rc != 0:
        raise GftError(rc, "gft_bench_copy failed")
    return ms.value


class Evolver:
    """One lattice on the GPU behind the C ABI."""

    def __init__(self, desc: GftDesc, _handle=None):
        self.desc = desc
        self.ns = desc.num_scalar if desc.model == 2 else 8 if desc.model == 1 else 4
        self.nv = link_reals(desc)
        self.n = (desc.nx, desc.ny, desc.nz)
        self.nsites = desc.nx * desc.ny * desc.nz
        self._owned = _handle is None
        if _handle is not None:
            self._h = _handle
            return
        h = C.c_void_p()
        rc = _L.gft_create(C.byref(desc), C.byref(h))
        if rc != 0:
            raise GftError(rc, _L.gft_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise GftError(rc, _L.gft_last_error(self._h).decode())

    def close(self):
        if self._h and self._owned:
            _L.gft_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def import_two_slot(self, scalar: np.ndarray, vec: np.ndarray, time_iter: int = 0):
        dt = scalar.dtype
        assert dt in (np.float32, np.float64) and vec.dtype == dt
        assert scalar.size == 2 * self.nsites * self.ns and vec.size == 2 * self.nsites * self.nv
        s = np.ascontiguousarray(scalar)
        v = np.ascontiguousarray(vec)
        T = C.c_float if dt == np.float32 else C.c_double
        fn = _L.gft_import_two_slot_f32 if dt == np.float32 else _L.gft_import_two_slot_f64
        self._check(fn(self._h, s.ctypes.data_as(C.POINTER(T)), v.ctypes.data_as(C.POINTER(T)), time_iter))

    def init_grf(self, seed: int, rms: float = 0.1, slope: float = -1.0, kmax: float = 0.0, vevs=(1.0, 1.0)):
        """Gaussian-random-field state generated on the GPU (gft_init_grf): time 0,
        both slots equal, identity links, E = 0."""
        g = GftGrfDesc(seed, rms, slope, kmax, (C.c_double * 2)(*vevs))
        self._check(_L.gft_init_grf(self._h, C.byref(g)))

    def init_gaussian(self, seed: int, sigma: float = 0.05, vevs=(1.0, 1.0)):
        """iid Gaussian state generated on the GPU (gft_init_gaussian): vev plus
        sigma per real component, time 0, both slots equal, identity links, E = 0."""
        g = GftGaussDesc(seed, sigma, (C.c_double * 2)(*vevs))
        self._check(_L.gft_init_gaussian(self._h, C.byref(g)))

    def export_two_slot(self, dtype=np.float64):
        s = np.zeros((2, self.nsites, self.ns), dtype)
        v = np.zeros((2, self.nsites, self.nv), dtype)
        T = C.c_float if dtype == np.float32 else C.c_double
        fn = _L.gft_export_two_slot_f32 if dtype == np.float32 else _L.gft_export_two_slot_f64
        it = C.c_uint(0)
        self._check(fn(self._h, s.ctypes.data_as(C.POINTER(T)), v.ctypes.data_as(C.POINTER(T)), C.byref(it)))
        return s, v, it.value

    def evolve(self, nsteps: int, energy_freq: int = 0, gauss_freq: int = 0, max_rows: int | None = None):
        if max_rows is None:
            max_rows = nsteps if (energy_freq or gauss_freq) else 0
        rows = (GftObs * max(max_rows, 1))()
        nrows = C.c_uint(0)
        self._check(_L.gft_evolve(self._h, nsteps, energy_freq, gauss_freq, rows, max_rows, C.byref(nrows)))
        return list(rows)[:nrows.value]

    def checkpoint(self) -> bytes:
        """The complete state (gft_checkpoint) as bytes."""
        n = C.c_size_t()
        self._check(_L.gft_checkpoint_size(self._h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._check(_L.gft_checkpoint(self._h, buf, n.value))
        return buf.raw

    def restore(self, data: bytes):
        self._check(_L.gft_restore(self._h, C.c_char_p(data), len(data)))

    def snapshot(self):
        self._check(_L.gft_snapshot(self._h))

    def rollback(self) -> int:
        it = C.c_uint()
        self._check(_L.gft_rollback(self._h, C.byref(it)))
        return it.value

    def observe(self) -> GftObs:
        o = GftObs()
        self._check(_L.gft_observe(self._h, C.byref(o)))
        return o

    def observe_noalloc(self) -> GftObs:
        """gft_observe_noalloc: the same observables without device allocation."""
        o = GftObs()
        self._check(_L.gft_observe_noalloc(self._h, C.byref(o)))
        return o

    def wilson_loops(self) -> GftLoops:
        o = GftLoops()
        self._check(_L.gft_wilson_loops(self._h, C.byref(o)))
        return o

    def set_launch(self, bz: int, by: int, order: int = 3, min_waves: int = 4, stage: int = 2):
        self._check(_L.gft_set_launch(self._h, bz, by, int(order), min_waves, int(stage)))

    def sync(self):
        self._check(_L.gft_sync(self._h))

    def bench_streams(self, reps: int = 20) -> float:
        """Average ms of streaming exactly the arrays one step reads and writes
        (gft_bench_streams); overwrites the state."""
        ms = C.c_double()
        self._check(_L.gft_bench_streams(self._h, reps, C.byref(ms)))
        return ms.value

    def last_step_ms(self):
        avg, tot, n = C.c_double(), C.c_double(), C.c_uint()
        self._check(_L.gft_last_step_ms(self._h, C.byref(avg), C.byref(tot), C.byref(n)))
        return avg.value, tot.value, n.value


class Ensemble:
    """Independent FP32 momentum-form lattices advanced together by one
    kernel launch per step (gft_ensemble_* in include/gft.h).  members[i]
    is an Evolver view of member i for import, export and observables; it
    stays valid while the ensemble is open."""

    def __init__(self, descs):
        arr = (GftDesc * len(descs))(*descs)
        h = C.c_void_p()
        rc = _L.gft_ensemble_create(arr, len(descs), C.byref(h))
        if rc != 0:
            raise GftError(rc, _L.gft_ensemble_last_error(None).decode())
        self._h = h
        self.members = []
        for i, d in enumerate(descs):
            m = C.c_void_p()
            self._check(_L.gft_ensemble_member(self._h, i, C.byref(m)))
            self.members.append(Evolver(d, _handle=m))

    def _check(self, rc):
        if rc != 0:
            raise GftError(rc, _L.gft_ensemble_last_error(self._h).decode())

    def close(self):
        if self._h:
            for m in self.members:
                m._h = None
            _L.gft_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_launch(self, member: int, bz: int, by: int):
        self._check(_L.gft_ensemble_set_launch(self._h, member, bz, by))

    def evolve(self, nsteps: int, use_graph: bool = True):
        self._check(_L.gft_ensemble_evolve(self._h, nsteps, int(use_graph)))

    def last_step_ms(self) -> float:
        ms = C.c_double()
        self._check(_L.gft_ensemble_last_step_ms(self._h, C.byref(ms)))
        return ms.value


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The surrogate network on the GPU (gft_mlp_* in include/gft.h)."""

    def __init__(self, n_in=8, width=256, n_hidden=4, n_heads=4, n_out=256, max_batch=512):
        self.desc = GftMlpDesc(n_in, width, n_hidden, n_heads, n_out, max_batch)
        h = C.c_void_p()
        rc = _L.gft_mlp_create(C.byref(self.desc), C.byref(h))
        if rc != 0:
            raise GftError(rc, _L.gft_mlp_last_error(None).decode())
        self._h = h
        self.n_params = int(_L.gft_mlp_num_params(h))

    def _check(self, rc):
        if rc != 0:
            raise GftError(rc, _L.gft_mlp_last_error(self._h).decode())

    def close(self):
        if self._h:
            _L.gft_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def layout(self):
        """[(name, shape)] of the flat parameter vector, in order."""
        d = self.desc
        out = []
        for l in range(d.n_hidden):
            out += [(f"W{l}", (d.width, d.n_in if l == 0 else d.width)), (f"b{l}", (d.width,))]
        out += [("Wh", (d.n_heads * d.n_out, d.width)), ("bh", (d.n_heads * d.n_out,))]
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(_L.gft_mlp_set_params(self._h, _fp(p)))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._check(_L.gft_mlp_get_params(self._h, _fp(p)))
        return p

    def forward(self, x: np.ndarray) -> np.ndarray:
        """[n, n_heads, n_out] for x [n, n_in]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        y = np.zeros((n, self.desc.n_heads, self.desc.n_out), np.float32)
        self._check(_L.gft_mlp_forward(self._h, _fp(x), n, _fp(y)))
        return y

    def gradients(self, x, t, w):
        x, t, w = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, w))
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(_L.gft_mlp_gradients(self._h, _fp(x), _fp(t), _fp(w), x.shape[0], _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x, t, w, batch=256, epochs=100, lr=1e-3, seed=0):
        x, t, w = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, w))
        hist = np.zeros(max(epochs, 1), np.float32)
        self._check(_L.gft_mlp_train(self._h, _fp(x), _fp(t), _fp(w), x.shape[0], batch, epochs, lr, seed, _fp(hist)))
        return hist[:epochs]


__all__ = ["Evolver", "Ensemble", "Mlp", "GftError", "device_info", "lib", "GFT_FORM_TWO_SLOT_F64", "GFT_FORM_TWO_SLOT_F32",
           "GFT_FORM_MOMENTUM_F32", "GFT_FORM_MOMENTUM_F64"]
