"""gft_observe_noalloc against gft_observe through the C ABI.

The allocation-free entry point runs the same reduction kernels with their
partials in the library's static scratch array: its observables must equal
gft_observe's bitwise, on two calls in a row (the second on a smaller lattice,
so the array holds the larger call's leftovers), and a lattice above the
array's capacity is refused before any launch."""
from __future__ import annotations

import numpy as np
import pytest

from gftevolver_amd._abi import GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64, GFT_EINVAL, make_desc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from gftevolver_amd import runtime
    return runtime


def _obs_array(o):
    return np.array(list(o.energy) + list(o.gauss_int[:4]) + list(o.gauss_max[:4]) + list(o.gauss_scale[:4]))


@pytest.mark.parametrize("form", [GFT_FORM_MOMENTUM_F32, GFT_FORM_MOMENTUM_F64], ids=["f32", "f64"])
def test_noalloc_observe_equals_observe(rt, form):
    for n, steps in (((48, 32, 64), 7), ((16, 12, 18), 3)):   # second call smaller
        with rt.Evolver(make_desc(form=form, n=n)) as ev:
            ev.init_grf(seed=5, rms=0.15)
            o0a = ev.observe_noalloc()      # before any step: energy only
            o0b = ev.observe()
            assert not o0a.has_gauss and not o0b.has_gauss
            np.testing.assert_array_equal(_obs_array(o0a), _obs_array(o0b))
            ev.evolve(steps)
            a = ev.observe_noalloc()
            b = ev.observe()
            c = ev.observe_noalloc()
            assert a.has_energy and a.has_gauss and a.time_iter == b.time_iter == steps
            print(f"{n}: energy {a.energy[0]:.12e}, Gauss int {a.gauss_int[0]:.3e}")
            np.testing.assert_array_equal(_obs_array(a), _obs_array(b))
            np.testing.assert_array_equal(_obs_array(c), _obs_array(b))
            assert np.isfinite(_obs_array(a)).all() and a.energy[0] > 0


def test_noalloc_refuses_lattice_above_capacity(rt):
    n = (256, 256, 344)   # 22 544 384 sites > 22 369 536
    with rt.Evolver(make_desc(form=GFT_FORM_MOMENTUM_F32, n=n)) as ev:
        with pytest.raises(rt.GftError) as e:
            ev.observe_noalloc()
        assert e.value.code == GFT_EINVAL
        assert "scratch" in str(e.value)
