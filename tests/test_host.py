"""Host-side logic of libgdp_hip (no device needed): the library loads and
exports every entry point of include/gdp_hip.h; edge numbering, DOF numbering,
CSR pattern and RCMK match the oracle bit for bit; solver.par handling;
device entry points fail loudly without a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from getdp_amd import _host, problems as P
from getdp_amd._lib import LIB_PATH, SIGNATURES, GdpError, check, lib
from dirichlet import with_dirichlet

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    txt = (ROOT / "include" / "gdp_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gdp_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert LIB_PATH.exists()
    L = lib()
    names = header_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    assert sorted(n for n, _, _ in SIGNATURES) == names
    assert L.missing_symbols == []


def test_device_calls_fail_loudly_without_gpu():
    import getdp_amd.linalg as LA
    try:
        import torch  # noqa: F401  (only to know whether a GPU is visible)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("GPU present")
    with pytest.raises(GdpError):
        LA.init(0)


@pytest.mark.parametrize("make", [lambda: P.curlcurl3d(2), lambda: P.curlcurl3d(3, seed=4)])
def test_edge_numbering_matches_oracle(oracle, make):
    pr = make()
    visit = pr.meta["visit"]
    e1, n1 = _host.number_edges(pr.etype, pr.conn, visit)
    e2, n2 = oracle.number_edges(pr.etype, pr.conn, visit)
    assert np.array_equal(e1, e2) and np.array_equal(n1, n2)
    n = pr.meta["n"]
    assert n1.shape[0] == 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    # sign +1 iff the local edge runs from the lower to the higher node number
    tets = pr.conn[: pr.meta["ntet"]]
    loc = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for k, (p, q) in enumerate(loc):
        s = np.sign(e1[: pr.meta["ntet"], k])
        assert np.array_equal(s, np.where(tets[:, p] < tets[:, q], 1, -1))


PROBLEMS = {
    "poisson2d_16": lambda: P.poisson2d(16),
    "poisson2d_lognormal_12": lambda: P.poisson2d(12, field_kind="lognormal", nbins=8),
    "elasticity3d_3": lambda: P.elasticity3d(3, nbins=8),
    "curlcurl3d_3": lambda: P.curlcurl3d(3),
    "poisson2d_az_12": lambda: P.poisson2d(12, basis="perp"),
    "poisson2d_dirichlet_dup": lambda: with_dirichlet(P.poisson2d(10), seed=1),
    "elasticity3d_dirichlet_dup": lambda: with_dirichlet(P.elasticity3d(3, nbins=8), seed=2),
    "curlcurl3d_dirichlet_dup": lambda: with_dirichlet(P.curlcurl3d(3), seed=3),
}


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_dof_numbering_and_pattern_match_oracle(oracle, name):
    from getdp_amd.linalg import FESystem
    pr = PROBLEMS[name]()
    ia, ja, a, b, tab = oracle.assemble(pr)
    fe = FESystem(pr)
    assert fe.N == ia.shape[0] - 1 and fe.nnz == ja.shape[0]
    for s in range(len(pr.spaces)):
        assert np.array_equal(fe.dof_table(s), tab[s])
    pia, pja = fe.pattern()
    assert np.array_equal(pia, ia) and np.array_equal(pja, ja)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_rcmk_matches_oracle(oracle, name):
    pr = PROBLEMS[name]()
    ia, ja, a, b, tab = oracle.assemble(pr)
    perm, rperm, nlev = _host.rcmk(ia, ja)
    o = oracle.rcmk(ia, ja, a)
    assert np.array_equal(perm, o[0]) and np.array_equal(rperm, o[1]) and nlev == o[5]


def test_unsupported_term_is_rejected():
    from getdp_amd.linalg import FESystem
    pr = P.poisson2d(4)
    pr.terms[0].kind = 77
    with pytest.raises(GdpError) as e:
        FESystem(pr)
    assert e.value.code == -4


def test_perpendicular_edge_limits_are_rejected():
    """BF_PerpendicularEdge is a 2D form: grad-grad on it, a nodal DOF side,
    tetrahedra and Jacobian Sur are refused (the reference's CPU path keeps
    them)."""
    from getdp_amd.linalg import FESystem
    pr = P.poisson2d(4, basis="perp")
    pr.terms[0].kind = P.T_GRAD_GRAD
    with pytest.raises(GdpError) as e:
        FESystem(pr)
    assert e.value.code == -4
    pr = P.poisson2d(4, basis="perp")
    pr.terms[0].jac_sur = 1
    with pytest.raises(GdpError) as e:
        FESystem(pr)
    assert e.value.code == -4
    pr = P.elasticity3d(1, nbins=2)
    for s in pr.spaces:
        s.bf = P.BF_PERP_EDGE
    for t in pr.terms:
        if t.kind == P.T_GRAD_GRAD:
            t.kind = P.T_CURL_CURL
    with pytest.raises(GdpError) as e:
        FESystem(pr)
    assert e.value.code == -4


def test_zero_row_is_reported():
    from getdp_amd.linalg import FESystem
    pr = P.poisson2d(4)
    pr.terms = pr.terms[1:]           # only the source term: no matrix entries
    with pytest.raises(GdpError) as e:
        FESystem(pr)
    assert e.value.code == -6


def test_solver_parameters(tmp_path):
    from getdp_amd.linalg import Solver
    s = Solver()
    # defaults of the reference's table (Sparskit.cpp:1847-1868)
    assert s.get("Preconditioner") == 2 and s.get("Algorithm") == 8
    assert s.get("Krylov_Size") == 40 and s.get("Nb_Iter_Max") == 1000
    assert s.get("Stopping_Test") == 1e-10 and s.get("Renumbering_Technique") == 1
    # per-option override rounds reals through float (init_solver_option)
    s.set("Stopping_Test", "1e-10")
    assert s.get("Stopping_Test") == float(np.float32(1e-10))
    with pytest.raises(GdpError):
        s.set("No_Such_Parameter", "1")
    f = tmp_path / "solver.par"
    f.write_text("/* comment */\n  Preconditioner 6\nKrylov_Size 30\n Stopping_Test 1e-12\n")
    s2 = Solver(str(f))
    assert s2.get("Preconditioner") == 6 and s2.get("Krylov_Size") == 30
    assert s2.get("Stopping_Test") == 1e-12


def test_batch_concatenation_is_block_diagonal():
    """The concatenated formulation numbers each instance's DOFs
    consecutively and couples no two instances (host pre-pass only)."""
    from getdp_amd import batch as BT
    from getdp_amd import linalg as LA
    probs = [P.poisson2d(n, seed=i, nblocks=4) for i, n in enumerate([5, 9, 7])]
    pr = BT.concat(probs)
    fe = LA.FESystem(pr)
    assert fe.N == sum((n - 1) ** 2 for n in [5, 9, 7])
    ia, ja = fe.pattern()
    off = np.cumsum([0] + [(n - 1) ** 2 for n in [5, 9, 7]])
    inst = np.searchsorted(off, np.arange(fe.N), side="right") - 1
    rows = np.repeat(np.arange(fe.N), np.diff(ia))
    assert np.array_equal(inst[rows], inst[ja])


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message."""
    import json
    import subprocess
    import sys
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"] and msg["n_gpus"] == 2
