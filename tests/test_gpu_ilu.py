"""Device ILU(0) application (factors + both sweeps) against the oracle's
restatement of ilu0 + lusol (ilut.F:1818-1941, :2297-2347), itself pinned bit
for bit to the reference's Fortran (tests/test_oracle_sparskit.py).

The device sweeps use one lane per row when rows have under 1.5 off-diagonal
entries on a side, else groups of 2-8 lanes (up to 32 on request) that wait
for several dependencies at once; all keep lusol's order of operations, so the
result must be bit-identical to the oracle's for FP32 and FP64 factor storage."""
import numpy as np
import pytest

from getdp_amd import problems as P

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "poisson2d_32": lambda: P.poisson2d(32),        # ~3 lower entries per row: 4 lanes
    "elasticity3d_4": lambda: P.elasticity3d(4, nbins=16),
    "elasticity3d_8": lambda: P.elasticity3d(8),    # ~18 lower entries: 8 lanes, 2-5 chunks a row
    "curlcurl3d_4": lambda: P.curlcurl3d(4),
    "poisson2d_192": lambda: P.poisson2d(192),  # 36,481 rows: wraps the one-wave sweeps' LDS ring
}


def _device_apply(gpu, ia, ja, a, v, f64, reps=1):
    n = ia.shape[0] - 1
    A = gpu.Matrix(n)
    A.set_csr(ia, ja, a)
    b = gpu.Vector(n, np.ones(n))
    x = gpu.Vector(n)
    s = gpu.Solver(Algorithm=8, Preconditioner=6, Krylov_Size=30, Renumbering_Technique=1,
                   ILU_Storage=1 if f64 else 0, Nb_Iter_Max=3)
    info = gpu.solve(A, b, s, x)  # builds the renumbering and the factors
    out = gpu.Vector(n)
    gpu.solver_prec_apply(s, gpu.Vector(n, v), out, reps=reps)
    return out.get(), info["ilu_sweep"]


SWEEP_KIND = {"free": 1, "wave": 2}


def _longest_side(ia, ja):
    """Most strictly-lower or strictly-upper entries in one row."""
    n = ia.shape[0] - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    lower = np.bincount(rows[ja < rows], minlength=n)
    upper = np.bincount(rows[ja > rows], minlength=n)
    return int(max(lower.max(), upper.max()))


@pytest.mark.parametrize("sweep", ["free", "wave", "wave-solo", "wave-team2"])
@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("name", list(SYSTEMS))
def test_ilu_apply_bit_exact(gpu, oracle, name, f64, sweep, monkeypatch):
    """Both sweep kinds: the sync-free sweeps and the level walk in one
    workgroup (solution entries in an LDS ring of 16384 rows; the 192^2
    system has 36,481 rows, so the ring wraps), walked by a team of waves
    (the default 4, and 2) or by one wave."""
    monkeypatch.setenv("GDP_ILU_SWEEP", sweep.split("-")[0])
    if sweep == "wave-solo":
        monkeypatch.setenv("GDP_ILU_TEAM", "1")
    elif sweep == "wave-team2":
        monkeypatch.setenv("GDP_ILU_TEAM", "2")
    print(f"case {name} f64={f64} sweep={sweep}", flush=True)
    ia, ja, a, b, tab = oracle.assemble(SYSTEMS[name]())
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    alu, jlu, ju, ierr = oracle.ilu0(ib, jb, bb, double=f64)
    assert ierr == 0
    v = np.random.default_rng(5).standard_normal(ib.shape[0] - 1)
    want = oracle.lusol(alu, jlu, ju, v)
    if sweep.startswith("wave") and _longest_side(ib, jb) > 32:
        # the level walk takes rows of up to 32 entries a side; forced on
        # longer ones it refuses instead of falling back
        from getdp_amd._lib import GdpError
        with pytest.raises(GdpError, match="entries on one side"):
            _device_apply(gpu, ia, ja, a, v, f64)
        return
    got, kind = _device_apply(gpu, ia, ja, a, v, f64, reps=3)  # repeated: sweep state reset each time
    assert kind == SWEEP_KIND[sweep.split("-")[0]], kind  # the forced kernel really ran
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), np.abs(got - want).max()


@pytest.mark.parametrize("f64", [False, True])
def test_ilu_apply_one_lane_rows(gpu, oracle, f64):
    """A tridiagonal system (one entry per side): the one-lane sweeps; with
    no fill in ILU(0) the application is the exact solve."""
    n = 3000
    rng = np.random.default_rng(11)
    ia = np.zeros(n + 1, np.int32)
    ja, a = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                ja.append(j)
                a.append(4.0 + rng.random() if j == i else -1.0 - rng.random())
        ia[i + 1] = len(ja)
    ja = np.array(ja, np.int32)
    a = np.array(a)
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    alu, jlu, ju, ierr = oracle.ilu0(ib, jb, bb, double=f64)
    v = rng.standard_normal(n)
    want = oracle.lusol(alu, jlu, ju, v)
    got, _ = _device_apply(gpu, ia, ja, a, v, f64)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    if f64:
        Bv = oracle.amux(ib, jb, bb, got)
        assert np.allclose(Bv, v, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("lanes", ["1", "2", "16", "32", "thread-rows"])
def test_ilu_apply_lane_counts(gpu, oracle, monkeypatch, lanes):
    """Every lanes-per-row variant of the sweeps (GDP_ILU_LANES overrides the
    choice made from the pattern) and the thread-per-row factorisation (used
    for rows beyond 1024 entries; GDP_ILU_THREAD_ROWS) give lusol's bits."""
    if lanes == "thread-rows":
        monkeypatch.setenv("GDP_ILU_THREAD_ROWS", "1")
    else:
        monkeypatch.setenv("GDP_ILU_LANES", lanes)
    ia, ja, a, b, tab = oracle.assemble(SYSTEMS["elasticity3d_8"]())
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    alu, jlu, ju, ierr = oracle.ilu0(ib, jb, bb)
    v = np.random.default_rng(6).standard_normal(ib.shape[0] - 1)
    want = oracle.lusol(alu, jlu, ju, v)
    got, kind = _device_apply(gpu, ia, ja, a, v, False)
    if lanes != "thread-rows":
        assert kind == SWEEP_KIND["free"]  # a lane count selects the sync-free sweeps
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("lanes", ["auto", "8", "wave"])
@pytest.mark.parametrize("pivot", [False, True])
@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("name", ["poisson2d_32", "elasticity3d_8", "curlcurl3d_4"])
def test_ilut_apply_bit_exact(gpu, oracle, name, f64, pivot, lanes, monkeypatch):
    """ILUT / ILUTP (fill 20, the reference's defaults) factored by the
    library on the host and applied by the device sweeps, against the oracle's
    ilut/ilutp + lusol; ILUTP's column permutation applied to the output."""
    if lanes == "wave":
        monkeypatch.setenv("GDP_ILU_SWEEP", "wave")
    elif lanes != "auto":
        monkeypatch.setenv("GDP_ILU_LANES", lanes)
    print(f"case {name} f64={f64} pivot={pivot} lanes={lanes}", flush=True)
    ia, ja, a, b, tab = oracle.assemble(SYSTEMS[name]())
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    n = ib.shape[0] - 1
    alu, jlu, ju, iperm, _, ierr = oracle.ilut(ib, jb, bb, 20, 0.0, 0.05, pivot, f64)
    assert ierr == 0
    v = np.random.default_rng(5).standard_normal(n)
    want = oracle.lusol(alu, jlu, ju, v)[iperm[n:]] if pivot else oracle.lusol(alu, jlu, ju, v)
    A = gpu.Matrix(n)
    A.set_csr(ia, ja, a)
    s = gpu.Solver(Algorithm=8, Preconditioner=2 if pivot else 1, Krylov_Size=40, ILU_Storage=1 if f64 else 0,
                   Nb_Iter_Max=3)
    info = gpu.solve(A, gpu.Vector(n, np.ones(n)), s, gpu.Vector(n))
    if lanes == "wave":
        assert info["ilu_sweep"] == SWEEP_KIND["wave"]
    elif lanes == "8":
        assert info["ilu_sweep"] == SWEEP_KIND["free"]
    out = gpu.Vector(n)
    gpu.solver_prec_apply(s, gpu.Vector(n, v), out, reps=3)
    got = out.get()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), np.abs(got - want).max()


@pytest.mark.parametrize("sweep", ["free", "wave", "wave-solo"])
@pytest.mark.parametrize("pivot", [False, True])
def test_ilut_apply_long_chain(gpu, oracle, sweep, pivot, monkeypatch):
    """ILUT / ILUTP(20) on Poisson 160^2 (25,281 rows, more than the one-wave
    sweeps' LDS ring holds, and thousands of narrow levels): lusol's bits with
    either sweep kind, a team of waves or one wave."""
    monkeypatch.setenv("GDP_ILU_SWEEP", sweep.split("-")[0])
    if sweep == "wave-solo":
        monkeypatch.setenv("GDP_ILU_TEAM", "1")
    ia, ja, a, b, tab = oracle.assemble(P.poisson2d(160))
    perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
    n = ib.shape[0] - 1
    alu, jlu, ju, iperm, _, ierr = oracle.ilut(ib, jb, bb, 20, 0.0, 0.05, pivot, False)
    assert ierr == 0
    v = np.random.default_rng(9).standard_normal(n)
    want = oracle.lusol(alu, jlu, ju, v)[iperm[n:]] if pivot else oracle.lusol(alu, jlu, ju, v)
    A = gpu.Matrix(n)
    A.set_csr(ia, ja, a)
    s = gpu.Solver(Algorithm=8, Preconditioner=2 if pivot else 1, Krylov_Size=40, Nb_Iter_Max=3)
    info = gpu.solve(A, gpu.Vector(n, np.ones(n)), s, gpu.Vector(n))
    assert info["ilu_sweep"] == SWEEP_KIND[sweep.split("-")[0]]
    out = gpu.Vector(n)
    gpu.solver_prec_apply(s, gpu.Vector(n, v), out, reps=2)
    got = out.get()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), np.abs(got - want).max()


def test_forced_wave_sweep_rejects_what_it_cannot_take(gpu, monkeypatch):
    """GDP_ILU_SWEEP=wave on a system whose dependencies leave the level
    walk's 16,384-entry LDS ring must fail with GDP_E_UNSUPPORTED, not fall
    back silently; unforced, the same system solves with the sync-free
    sweeps.  The system: 40,000 rows, each coupled to its neighbours (one
    row per level) and to the rows 20,000 away, so a dependency was solved
    20,000 stream positions earlier."""
    from getdp_amd._lib import GdpError
    n, far = 40_000, 20_000
    ia, ja, a = [0], [], []
    for i in range(n):
        for j in (i - far, i - 1, i, i + 1, i + far):
            if 0 <= j < n:
                ja.append(j)
                a.append(6.0 if j == i else -1.0)
        ia.append(len(ja))
    ia, ja, a = np.array(ia, np.int32), np.array(ja, np.int32), np.array(a)
    A = gpu.Matrix(n)
    A.set_csr(ia, ja, a)
    opts = dict(Algorithm=8, Preconditioner=6, Krylov_Size=30, Renumbering_Technique=0, Nb_Iter_Max=2)
    monkeypatch.setenv("GDP_ILU_SWEEP", "wave")
    with pytest.raises(GdpError) as e:
        gpu.solve(A, gpu.Vector(n, np.ones(n)), gpu.Solver(**opts), gpu.Vector(n))
    assert "ring" in str(e.value)
    monkeypatch.delenv("GDP_ILU_SWEEP")
    info = gpu.solve(A, gpu.Vector(n, np.ones(n)), gpu.Solver(**opts), gpu.Vector(n))
    assert info["ilu_sweep"] == SWEEP_KIND["free"]
