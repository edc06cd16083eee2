"""GPU test of the allocation-free FP32 direct sum (nb_force_direct_f32_noalloc):
called through the C ABI, it gives the same bits as an NB_DIRECT_F32
calculator (nb_force_compute) on the same particles, on consecutive calls
whose second input is smaller than the first (the shared scratch array keeps
the first call's contents), with and without softening; sizes beyond its
scratch capacity are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def reference_f32(nb, d, eps, G):
    f = nb.ForceCalculatorNaive(eps, "f32", G=G)
    p = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"])
    f.computeAccelerations(p)
    return p.acc()


def noalloc(nb, d, eps, G):
    p = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"])
    rc = nb.lib.nb_force_direct_f32_noalloc(p.ref, C.c_double(G), C.c_double(eps), None)
    assert rc == 0, nb.lib.nb_last_error().decode()
    return p.acc()


@pytest.mark.parametrize("eps", [0.01, 0.0])
def test_noalloc_equals_calculator_on_consecutive_calls(nb, eps):
    big = nb.ic_plummer(70001, 11)
    small = nb.ic_plummer(5003, 12)
    for d in (big, small):  # second call smaller: stale scratch beyond n is never read
        a = noalloc(nb, d, eps, 1.0)
        r = reference_f32(nb, d, eps, 1.0)
        assert all(np.array_equal(x, y) for x, y in zip(a, r))
        assert all(np.isfinite(x).all() for x in a)


def test_noalloc_limits(nb):
    assert nb.NB_NOALLOC_MAX_N == 1 << 20
    p = nb.Particles(nb.NB_NOALLOC_MAX_N + 1)
    rc = nb.lib.nb_force_direct_f32_noalloc(p.ref, C.c_double(1.0), C.c_double(0.01), None)
    assert rc == 4 and b"scratch capacity" in nb.lib.nb_last_error()  # NB_ERR_CAPACITY
    empty = nb.Particles(0)
    assert nb.lib.nb_force_direct_f32_noalloc(empty.ref, C.c_double(1.0), C.c_double(0.0),
                                              None) == 0
    assert nb.lib.nb_force_direct_f32_noalloc(None, C.c_double(1.0), C.c_double(0.0), None) == 1
    # the largest size it accepts, against the calculator
    d = nb.ic_uniform_gaussian(nb.NB_NOALLOC_MAX_N, 5)
    q = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"])
    nb.direct_f32_noalloc(q, 0.01, G=1.0)
    r = reference_f32(nb, d, 0.01, 1.0)
    assert all(np.array_equal(x, y) for x, y in zip(q.acc(), r))
