"""The allocation-free dot product (gdp_vector_dot_noalloc) against
gdp_vector_dot through the C-ABI: same kernels and summation order, so the
results must be the same bits, on two calls in a row (the second on shorter
vectors, over the partials the first left in the shared array)."""
import ctypes as C

import numpy as np
import pytest

from getdp_amd._lib import lib

pytestmark = pytest.mark.gpu


def _dot(fn, U, V):
    r = C.c_double()
    rc = fn(U.h, V.h, C.byref(r))
    assert rc == 0, lib().gdp_last_error()
    return r.value


def test_dot_noalloc_matches_dot_twice(gpu):
    L = lib()
    rng = np.random.default_rng(21)
    for n in (2_000_003, 1_537):
        print(f"case n={n}", flush=True)
        u, v = rng.standard_normal(n), rng.standard_normal(n) * 1e2
        U, V = gpu.Vector(n, u), gpu.Vector(n, v)
        want = _dot(L.gdp_vector_dot, U, V)
        got = _dot(L.gdp_vector_dot_noalloc, U, V)
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (got, want)
        assert abs(got - float(np.dot(u, v))) <= n * 2.3e-16 * float(np.abs(u * v).sum())


def test_dot_noalloc_refusals(gpu):
    L = lib()
    U, V = gpu.Vector(10, np.ones(10)), gpu.Vector(11, np.ones(11))
    r = C.c_double()
    assert L.gdp_vector_dot_noalloc(U.h, V.h, C.byref(r)) != 0  # size mismatch
    W = gpu.Vector(10, np.ones(10))
    W.add_double(3, 1.0)
    assert L.gdp_vector_dot_noalloc(U.h, W.h, C.byref(r)) != 0  # pending additions
    assert b"pending" in lib().gdp_last_error()
    W.assemble()
    assert _dot(L.gdp_vector_dot_noalloc, U, W) == 11.0
    E = gpu.Vector(0)
    assert _dot(L.gdp_vector_dot_noalloc, E, E) == 0.0
