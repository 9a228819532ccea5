"""Whole-vector and whole-matrix LinAlg arithmetic on the device against the
reference's sequential loops (Sparskit.cpp:797-1215), restated here with
numpy's elementwise operations (one rounding per operation, like the C
loops): bit for bit for the elementwise operations and the inf-norm; the
dot product within the bound of a reordered sum, n eps sum|u_i v_i|; the
reference's argument rules (result aliasing) enforced."""
import math

import numpy as np
import pytest

from getdp_amd import problems as P
from getdp_amd._lib import GdpError, lib

pytestmark = pytest.mark.gpu


def _ok(rc):
    assert rc == 0, lib().gdp_last_error()


def _eq(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def test_vector_ops_match_sequential_loops(gpu):
    L = lib()
    n = 100_003
    rng = np.random.default_rng(4)
    u0, v0 = rng.standard_normal(n), rng.standard_normal(n) * 1e3
    d = -0.37
    U, V = gpu.Vector(n, u0), gpu.Vector(n, v0)
    _ok(L.gdp_vector_add(U.h, V.h, U.h))
    assert _eq(U.get(), u0 + v0)
    U.set(u0)
    _ok(L.gdp_vector_add_scaled(U.h, V.h, d, U.h))
    assert _eq(U.get(), u0 + d * v0)
    U.set(u0)
    _ok(L.gdp_vector_sub(U.h, V.h, U.h))
    assert _eq(U.get(), u0 - v0)
    U.set(u0)
    _ok(L.gdp_vector_sub(U.h, V.h, V.h))
    assert _eq(V.get(), u0 - v0) and _eq(U.get(), u0)
    V.set(v0)
    _ok(L.gdp_vector_scale(U.h, d, U.h))
    assert _eq(U.get(), u0 * d)
    U.set(u0)
    _ok(L.gdp_vector_swap(U.h, V.h))
    assert _eq(U.get(), v0) and _eq(V.get(), u0)
    _ok(L.gdp_vector_set_all(U.h, 2.5))
    assert np.all(U.get() == 2.5)
    # dot: deterministic, within n eps sum|u v| of the exact sum
    U.set(u0)
    V.set(v0)
    import ctypes as C
    r1, r2 = C.c_double(), C.c_double()
    _ok(L.gdp_vector_dot(U.h, V.h, C.byref(r1)))
    _ok(L.gdp_vector_dot(U.h, V.h, C.byref(r2)))
    exact = math.fsum((u0 * v0).tolist())
    assert r1.value == r2.value
    assert abs(r1.value - exact) <= n * 2.2e-16 * float(np.sum(np.abs(u0 * v0)))
    # inf-norm: exact, NaN entries ignored as the reference's comparison does
    w0 = u0.copy()
    w0[[5, 77_000]] = np.nan
    w0[123] = -40.0
    W = gpu.Vector(n, w0)
    _ok(L.gdp_vector_norm_inf(W.h, C.byref(r1)))
    assert r1.value == 40.0
    # dummy DOFs
    dummy = (rng.random(n) < 0.1).astype(np.int32)
    U.set(u0)
    _ok(L.gdp_vector_dummy(U.h, dummy.ctypes.data_as(C.POINTER(C.c_int))))
    assert _eq(U.get(), np.where(dummy == 1, 0.0, u0))
    # the reference's argument rules
    X = gpu.Vector(n)
    for call in (lambda: L.gdp_vector_add(U.h, V.h, X.h), lambda: L.gdp_vector_add_scaled(U.h, V.h, d, V.h),
                 lambda: L.gdp_vector_sub(U.h, V.h, X.h), lambda: L.gdp_vector_scale(U.h, d, V.h)):
        assert call() == -1 and "Wrong arguments" in L.gdp_last_error().decode()
    assert L.gdp_vector_swap(U.h, gpu.Vector(n - 1).h) == -1


def _matrix(gpu, ia, ja, a):
    n = ia.shape[0] - 1
    M = gpu.Matrix(n)
    M.set_csr(ia, ja, a)
    return M


def _dense(ia, ja, a, n):
    D = np.zeros((n, n))
    for i in range(n):
        D[i, ja[ia[i]:ia[i + 1]]] = a[ia[i]:ia[i + 1]]
    return D


def test_matrix_ops_same_and_different_patterns(gpu, oracle):
    import ctypes as C
    L = lib()
    ia, ja, a, b, tab = oracle.assemble(P.poisson2d(16))
    n = ia.shape[0] - 1
    rng = np.random.default_rng(2)
    a2 = rng.standard_normal(a.shape[0])
    d = 1.75
    A, B = _matrix(gpu, ia, ja, a), _matrix(gpu, ia, ja, a2)
    _ok(L.gdp_matrix_add_scaled(A.h, B.h, d, A.h))
    ia_, ja_, got = A.get_csr()
    assert _eq(got, a + d * a2)
    _ok(L.gdp_matrix_add(A.h, B.h, B.h))  # m3 == m2: m2 += m1
    assert _eq(B.get_csr()[2], a2 + (a + d * a2))
    _ok(L.gdp_matrix_scale(A.h, -3.0, A.h))
    assert _eq(A.get_csr()[2], (a + d * a2) * -3.0)
    # the operator follows the new values
    x = rng.standard_normal(n)
    X, Y = gpu.Vector(n, x), gpu.Vector(n)
    gpu.prod_matrix_vector(A, X, Y)
    assert _eq(Y.get(), oracle.amux(ia, ja, (a + d * a2) * -3.0, x))
    # different patterns: the union, entries of the second alone scaled by d
    C_ia = np.arange(n + 1, dtype=np.int32)
    C_ja = ((np.arange(n) + 7) % n).astype(np.int32)  # one entry per row, mostly outside A's pattern
    C_a = rng.standard_normal(n)
    A2 = _matrix(gpu, ia, ja, a)
    Cm = _matrix(gpu, C_ia, C_ja, C_a)
    _ok(L.gdp_matrix_add_scaled(A2.h, Cm.h, d, A2.h))
    ia2, ja2, a2v = A2.get_csr()
    want = _dense(ia, ja, a, n)
    mask = np.zeros((n, n), bool)
    mask[np.repeat(np.arange(n), np.diff(ia)), ja] = True
    for i in range(n):
        j = C_ja[i]
        want[i, j] = want[i, j] + d * C_a[i] if mask[i, j] else d * C_a[i]
        mask[i, j] = True
    assert ia2[-1] == mask.sum()
    for i in range(n):
        cols = ja2[ia2[i]:ia2[i + 1]]
        assert np.array_equal(cols, np.nonzero(mask[i])[0])
        assert _eq(a2v[ia2[i]:ia2[i + 1]], want[i, cols])
    nn, mm = C.c_int(), C.c_int()
    _ok(L.gdp_matrix_size(A2.h, C.byref(nn), C.byref(mm)))
    assert nn.value == n
    assert L.gdp_matrix_add(A2.h, Cm.h, _matrix(gpu, C_ia, C_ja, C_a).h) == -1


def test_staged_additions_count_on_every_operand(gpu):
    """LinAlg_AddDoubleInVector adds straight into the vector in the
    reference (LinAlg_SPARSKIT.cpp:356-364): pending additions must be seen
    by every operation that reads a vector, on either operand, and dropped by
    every operation that overwrites all of it."""
    import ctypes as C
    L = lib()
    n = 1000
    rng = np.random.default_rng(9)
    u0, v0 = rng.standard_normal(n), rng.standard_normal(n)
    adds = [(3, 0.5), (17, -2.0), (3, 1.25), (999, 4.0)]

    def staged(base):
        V = gpu.Vector(n, base)
        for i, d in adds:
            V.add_double(i, d)
        want = base.copy()
        for i, d in adds:
            want[i] = want[i] + d
        return V, want

    # second operand of add / scaled add / subtract (both forms) / swap
    for op in ("add", "add_scaled", "sub", "sub_rev", "swap"):
        U = gpu.Vector(n, u0)
        V, vw = staged(v0)
        if op == "add":
            _ok(L.gdp_vector_add(U.h, V.h, U.h))
            assert _eq(U.get(), u0 + vw)
        elif op == "add_scaled":
            _ok(L.gdp_vector_add_scaled(U.h, V.h, 0.3, U.h))
            assert _eq(U.get(), u0 + 0.3 * vw)
        elif op == "sub":
            _ok(L.gdp_vector_sub(U.h, V.h, U.h))
            assert _eq(U.get(), u0 - vw)
        elif op == "sub_rev":
            _ok(L.gdp_vector_sub(U.h, V.h, V.h))
            assert _eq(V.get(), u0 - vw)
        else:
            _ok(L.gdp_vector_swap(U.h, V.h))
            assert _eq(U.get(), vw) and _eq(V.get(), u0)
    # first operand too
    U, uw = staged(u0)
    V = gpu.Vector(n, v0)
    _ok(L.gdp_vector_add(U.h, V.h, U.h))
    assert _eq(U.get(), uw + v0)
    # dot, inf-norm, 2-norm, get
    U, uw = staged(u0)
    V, vw = staged(v0)
    r = C.c_double()
    _ok(L.gdp_vector_dot(U.h, V.h, C.byref(r)))
    assert abs(r.value - math.fsum((uw * vw).tolist())) <= n * 2.2e-16 * np.abs(uw * vw).sum()
    big = np.zeros(n)
    B = gpu.Vector(n, big)
    B.add_double(5, -7.0)
    _ok(L.gdp_vector_norm_inf(B.h, C.byref(r)))
    assert r.value == 7.0
    B = gpu.Vector(n, big)
    B.add_double(5, 3.0)
    B.add_double(6, 4.0)
    assert B.norm2() == 5.0
    U, uw = staged(u0)
    assert _eq(U.get(), uw)
    # set overwrites earlier additions: a later assemble must not re-add them
    U, uw = staged(u0)
    U.set(v0)
    U.assemble()
    assert _eq(U.get(), v0)
    U, uw = staged(u0)
    _ok(L.gdp_vector_set_all(U.h, 1.5))
    U.assemble()
    assert np.all(U.get() == 1.5)
