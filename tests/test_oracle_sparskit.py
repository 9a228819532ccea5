"""The oracle's Sparskit restatement against the reference's own kernels.

* Golden fixtures (tests/golden/sparskit_ref.npz) hold outputs of the
  reference's Fortran routines (cmkreord/sortcol, ilu0, solve_matrix's
  GMRES/FGMRES/CG/BCGSTAB/TFQMR sequence); the restatement must reproduce them bit for bit.
* Where oracle/_ref is built (this container), the restatement is also run
  side by side with the reference kernels on random sparse matrices.
"""
import numpy as np
import pytest

CASES = ["poisson2d_8", "elasticity3d_2", "curlcurl3d_2"]
SOLVES = {
    "gmres_ilu0": dict(algorithm=8, preconditioner=6, krylov=30),
    "fgmres_ilu0": dict(algorithm=9, preconditioner=6, krylov=30),
    "cg_jacobi": dict(algorithm=1, preconditioner=0, scaling=1),
    "gmres_none_natural": dict(algorithm=8, preconditioner=0, renumbering=0, krylov=10),
    "bcgstab_ilu0": dict(algorithm=5, preconditioner=6),
    "bcgstab_jacobi": dict(algorithm=5, preconditioner=0, scaling=1),
    "tfqmr_ilu0": dict(algorithm=6, preconditioner=6),
    "tfqmr_none": dict(algorithm=6, preconditioner=0),
    "gmres40_ilutp_default": dict(algorithm=8, preconditioner=2, krylov=40),
    "gmres40_ilut_fill5": dict(algorithm=8, preconditioner=1, krylov=40, nb_fill=5, droptol=1e-3),
}


def _sys(golden, name):
    return (golden[f"{name}/ia"], golden[f"{name}/ja"], golden[f"{name}/a"], golden[f"{name}/b"])


@pytest.mark.parametrize("name", CASES)
def test_rcmk_matches_reference(oracle, golden, name):
    ia, ja, a, b = _sys(golden, name)
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    assert np.array_equal(perm, golden[f"{name}/rcmk_perm"])
    assert np.array_equal(np.sort(perm), np.arange(perm.shape[0]))
    assert np.array_equal(rperm[perm], np.arange(perm.shape[0]))


@pytest.mark.parametrize("name", CASES)
def test_ilu0_f32_matches_reference(oracle, golden, name):
    ia, ja, a, b = _sys(golden, name)
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    alu, jlu, ju, ierr = oracle.ilu0(ib, jb, bb)
    ralu, rjlu, rju = (golden[f"{name}/ilu_alu"], golden[f"{name}/ilu_jlu"], golden[f"{name}/ilu_ju"])
    n = ib.shape[0] - 1
    used = int(jlu[n])  # first free slot
    assert ierr == 0
    assert np.array_equal(alu[:used].view(np.uint32), ralu[:used].view(np.uint32))
    assert np.array_equal(jlu[:used] + 1, rjlu[:used])
    assert np.array_equal(ju[:n] + 1, rju[:n])


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("sname", list(SOLVES))
def test_solve_matches_reference(oracle, golden, name, sname):
    ia, ja, a, b = _sys(golden, name)
    x, info, r = oracle.solve(ia, ja, a, b, maxits=20000, **SOLVES[sname])
    assert info["its"] == int(golden[f"{name}/{sname}/its"])
    assert info["status"] == int(golden[f"{name}/{sname}/status"])
    # identical arithmetic: bit-identical solutions
    assert np.array_equal(x, golden[f"{name}/{sname}/x"])
    # the reference reports dnrm2 (Lawson's scaled algorithm); the oracle a
    # plain sqrt of the sum of squares
    tr = float(golden[f"{name}/{sname}/true_res"])
    assert abs(info["true_res"] - tr) <= 1e-12 * max(tr, 1e-300) + 1e-300


def _random_csr(rng, n, extra, diag_boost=4.0):
    rows = [set([i]) for i in range(n)]
    for _ in range(extra):
        i, j = rng.integers(0, n, 2)
        rows[i].add(j)
        rows[j].add(i)
    ia = np.zeros(n + 1, np.int32)
    ja, a = [], []
    for i in range(n):
        cols = sorted(rows[i])
        ia[i + 1] = ia[i] + len(cols)
        for c in cols:
            ja.append(c)
            a.append(diag_boost * len(cols) if c == i else rng.standard_normal())
    return ia, np.array(ja, np.int32), np.array(a)


@pytest.mark.parametrize("seed", range(4))
def test_restatement_vs_reference_kernels_random(oracle, seed):
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(seed)
    ia, ja, a = _random_csr(rng, 200 + 37 * seed, 600)
    x = rng.standard_normal(ia.shape[0] - 1)
    assert np.array_equal(oracle.amux(ia, ja, a, x), oracle.ref_amux(ia, ja, a, x))
    p1 = oracle.rcmk(ia, ja, a)
    p2 = oracle.ref_rcmk(ia, ja, a)
    assert np.array_equal(p1[0], p2[0]) and np.array_equal(p1[3], p2[3])
    for dbl in (False, True):
        f1 = oracle.ilu0(p1[2], p1[3], p1[4], double=dbl)
        f2 = oracle.ref_ilu0(p1[2], p1[3], p1[4], double=dbl)
        n = ia.shape[0] - 1
        used = int(f1[1][n])
        assert np.array_equal(f1[0][:used], f2[0][:used])
        y = rng.standard_normal(n)
        assert np.array_equal(oracle.lusol(f1[0], f1[1], f1[2], y), oracle.ref_lusol(f2[0], f2[1], f2[2], y))
    b = rng.standard_normal(ia.shape[0] - 1)
    for kw in (dict(algorithm=8, preconditioner=6), dict(algorithm=9, preconditioner=6),
               dict(algorithm=8, preconditioner=0, krylov=5), dict(algorithm=1, preconditioner=0),
               dict(algorithm=5, preconditioner=6), dict(algorithm=5, preconditioner=0),
               dict(algorithm=6, preconditioner=6), dict(algorithm=6, preconditioner=0)):
        xo, io, _ = oracle.solve(ia, ja, a, b, maxits=3000, **kw)
        xr, ir, _ = oracle.ref_solve(ia, ja, a, b, maxits=3000, **kw)
        assert io["its"] == ir["its"] and io["status"] == ir["status"]
        assert np.array_equal(xo, xr)


def test_ilu0_is_exact_lu_on_tridiagonal(oracle):
    """ILU(0) of a tridiagonal matrix has no dropped fill: it is the exact LU
    factorisation, so one lusol solves the system (textbook identity)."""
    n = 50
    ia = np.zeros(n + 1, np.int32)
    ja, a = [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.5), (i + 1, -1.0)):
            if 0 <= j < n:
                ja.append(j)
                a.append(v)
        ia[i + 1] = len(ja)
    ja = np.array(ja, np.int32)
    a = np.array(a)
    alu, jlu, ju, ierr = oracle.ilu0(ia, ja, a, double=True)
    x = np.random.default_rng(0).standard_normal(n)
    y = oracle.amux(ia, ja, a, x)
    assert np.allclose(oracle.lusol(alu, jlu, ju, y), x, rtol=1e-12, atol=1e-12)


def test_gmres_full_krylov_solves_small_system_exactly(oracle):
    """GMRES(m) with m >= n terminates with the exact solution in at most n
    Arnoldi steps (exact-arithmetic identity, checked to round-off)."""
    rng = np.random.default_rng(3)
    n = 12
    M = rng.standard_normal((n, n)) + n * np.eye(n)
    ia = np.arange(0, n * n + 1, n, dtype=np.int32)
    ja = np.tile(np.arange(n, dtype=np.int32), n)
    b = rng.standard_normal(n)
    x, info, _ = oracle.solve(ia, ja, M.ravel(), b, algorithm=8, preconditioner=0, krylov=30,
                              renumbering=0, rtol=1e-14)
    assert info["its"] <= n + 1
    assert np.allclose(M @ x, b, rtol=0, atol=1e-11 * np.linalg.norm(b))


def test_zero_pivot_reported(oracle):
    ia = np.array([0, 2, 4], np.int32)
    ja = np.array([0, 1, 0, 1], np.int32)
    a = np.array([0.0, 1.0, 1.0, 1.0])
    alu, jlu, ju, ierr = oracle.ilu0(ia, ja, a)
    assert ierr == 1


@pytest.mark.parametrize("algorithm,its", [(5, 5), (6, 9)])
def test_short_iteration_limits_match_reference(oracle, algorithm, its):
    """BCGSTAB tests the limit after the second product of an iteration
    (4 -> 5 products), TFQMR at the end of an iteration (9 -> 9)."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(7)
    ia, ja, a = _random_csr(rng, 300, 900, diag_boost=1.0)
    b = rng.standard_normal(ia.shape[0] - 1)
    maxits = 4 if algorithm == 5 else 9
    xo, io, _ = oracle.solve(ia, ja, a, b, maxits=maxits, algorithm=algorithm, preconditioner=6)
    xr, ir, _ = oracle.ref_solve(ia, ja, a, b, maxits=maxits, algorithm=algorithm, preconditioner=6)
    assert io["status"] == ir["status"] == -1
    assert io["its"] == ir["its"] == its
    assert np.array_equal(xo, xr)


@pytest.mark.parametrize("algorithm,status", [(5, 0), (6, -3)])
def test_zero_rhs_matches_reference(oracle, golden, algorithm, status):
    """BCGSTAB stops before its first iteration on a zero residual; TFQMR has
    no such test and reports the breakdown rho = <r0, r> = 0 (-3)."""
    ia, ja, a, b = _sys(golden, "poisson2d_8")
    x, info, _ = oracle.solve(ia, ja, a, np.zeros_like(b), algorithm=algorithm, preconditioner=6)
    assert info["status"] == status and info["its"] == 1
    assert not np.any(x)
    if oracle.ref() is not None:
        xr, ir, _ = oracle.ref_solve(ia, ja, a, np.zeros_like(b), algorithm=algorithm, preconditioner=6)
        assert ir["status"] == status and ir["its"] == 1


def _ilut_cases(oracle, golden):
    rng = np.random.default_rng(21)
    out = []
    for boost in (1.0, 0.05):  # weak diagonals make ilutp permute columns
        ia, ja, a = _random_csr(rng, 180, 600, diag_boost=boost)
        out.append((ia, ja, a))
    for name in CASES:
        ia, ja, a, b = _sys(golden, name)
        perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
        out.append((ib, jb, bb))
    return out


@pytest.mark.parametrize("pivot", [False, True])
@pytest.mark.parametrize("double", [False, True])
def test_ilut_ilutp_match_reference(oracle, golden, pivot, double):
    """ilut / ilutp factors (values, columns, row pointers, permutation and
    the in-place column permutation of the matrix) bit for bit."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    permuted = 0
    for ia, ja, a in _ilut_cases(oracle, golden):
        n = ia.shape[0] - 1
        for lfil, droptol, permtol in ((20, 0.0, 0.05), (5, 1e-3, 0.05), (3, 0.0, 0.5)):
            o = oracle.ilut(ia, ja, a, lfil, droptol, permtol, pivot, double)
            r = oracle.ref_ilut(ia, ja, a, lfil, droptol, permtol, pivot, double)
            assert o[5] == r[5] == 0
            used = int(o[1][n])
            v = np.uint64 if double else np.uint32
            assert np.array_equal(o[0][:used].view(v), r[0][:used].view(v))
            assert np.array_equal(o[1][:used], r[1][:used])
            assert np.array_equal(o[2][:n], r[2][:n])
            if pivot:
                assert np.array_equal(o[3], r[3]) and np.array_equal(o[4], r[4])
                permuted += int(np.any(o[3][:n] != np.arange(n)))
    if pivot:
        assert permuted > 0  # the pivoting path was exercised


def test_ilutp_solve_with_pivoting_matches_reference(oracle):
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(5)
    ia, ja, a = _random_csr(rng, 150, 300, diag_boost=0.3)
    b = rng.standard_normal(150)
    for kw in (dict(preconditioner=2, permtol=0.5), dict(preconditioner=1)):
        xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=8, krylov=40, maxits=400, renumbering=0, **kw)
        xr, ir, _ = oracle.ref_solve(ia, ja, a, b, algorithm=8, krylov=40, maxits=400, renumbering=0, **kw)
        assert io["its"] == ir["its"] and io["status"] == ir["status"]
        assert np.array_equal(xo, xr)
