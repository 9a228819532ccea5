"""Device assembly and SpMV against the oracle (same inputs).

Contract: pattern and DOF numbering bit-exact; matrix entries within 1e-12
relative (the device path sums each entry in the reference's order and is
compiled without contraction, so the entries are checked bit for bit)."""
import numpy as np
import pytest

from getdp_amd import problems as P
from dirichlet import with_dirichlet

pytestmark = pytest.mark.gpu


def az_general(n=12, seed=6):
    """a_z (BF_PerpendicularEdge) on a jittered mesh: general triangles, an
    anisotropic tensor reluctivity per region and a mass term."""
    pr = P.poisson2d(n, seed=seed, field_kind="lognormal", sigma=1.0, nbins=8, basis="perp")
    rng = np.random.default_rng(seed)
    inner = np.all((pr.xyz[:, :2] > 0) & (pr.xyz[:, :2] < 1), axis=1)
    pr.xyz[inner, :2] += rng.uniform(-0.3 / n, 0.3 / n, size=(int(inner.sum()), 2))
    nu = pr.terms[0].coef[:, 0]
    C = np.zeros((nu.shape[0], 9))
    C[:, 0], C[:, 1], C[:, 3], C[:, 4], C[:, 8] = nu, 0.3 * nu, -0.2 * nu, 2.0 * nu, nu
    pr.terms[0].coef = C
    dom = pr.terms[0].regions
    pr.terms.append(P.Term(P.T_EDGE_MASS, 0, 0, dom, 3, np.full((C.shape[0], 1), 1e-2), dim=2))
    return pr


CASES = {
    "poisson2d_32": lambda: P.poisson2d(32),
    "poisson2d_lognormal_24": lambda: P.poisson2d(24, field_kind="lognormal", sigma=1.5, nbins=16),
    "poisson2d_1pt": lambda: P.poisson2d(8, npoints=1),
    "poisson2d_4pt": lambda: P.poisson2d(8, npoints=4),
    "elasticity3d_4": lambda: P.elasticity3d(4, nbins=16),
    "elasticity3d_1pt": lambda: P.elasticity3d(2, nbins=4, npoints=1),
    "curlcurl3d_4": lambda: P.curlcurl3d(4),
    "curlcurl3d_1pt": lambda: P.curlcurl3d(3, npoints=1),
    "poisson2d_az_32": lambda: P.poisson2d(32, basis="perp"),
    "poisson2d_az_general": az_general,
}


# non-zero fixed values, some entities listed twice with different values
# (the first listed value must win), on every family
DIRICHLET_CASES = {
    "poisson2d_24_dirichlet": lambda: with_dirichlet(P.poisson2d(24), seed=11),
    "elasticity3d_4_dirichlet": lambda: with_dirichlet(P.elasticity3d(4, nbins=16), seed=12),
    "curlcurl3d_4_dirichlet": lambda: with_dirichlet(P.curlcurl3d(4), seed=13),
    "poisson2d_az_16_dirichlet": lambda: with_dirichlet(P.poisson2d(16, basis="perp"), seed=14),
}
CASES.update(DIRICHLET_CASES)


@pytest.mark.parametrize("name", list(CASES))
def test_assembly_bit_exact(gpu, oracle, name):
    pr = CASES[name]()
    print(f"case {name}", flush=True)
    ia, ja, a, b, tab = oracle.assemble(pr)
    fe, A, bv = gpu.assemble(pr)
    gia, gja, ga = A.get_csr()
    assert np.array_equal(gia, ia) and np.array_equal(gja, ja)
    for s in range(len(pr.spaces)):
        assert np.array_equal(fe.dof_table(s), tab[s])
    rel = np.abs(ga - a).max() / np.abs(a).max()
    assert rel <= 1e-12
    assert np.array_equal(ga, a), f"max rel diff {rel}"
    assert np.array_equal(bv.get(), b)


def test_generate_twice_and_rebind(gpu, oracle):
    """Generate[A] zeroes and re-fills values; a second matrix adopts the
    pattern (SolvingOperations.cpp:175-184 recreates A on each Generate)."""
    pr = P.poisson2d(16)
    ia, ja, a, b, tab = oracle.assemble(pr)
    fe, A, bv = gpu.assemble(pr)
    fe.generate(A, bv)
    assert np.array_equal(A.get_csr()[2], a) and np.array_equal(bv.get(), b)
    A2 = gpu.Matrix(fe.N)
    b2 = gpu.Vector(fe.N)
    fe.generate(A2, b2)
    assert np.array_equal(A2.get_csr()[2], a)


@pytest.mark.parametrize("name", ["poisson2d_32", "elasticity3d_4", "curlcurl3d_4"])
def test_spmv_matches_amux(gpu, oracle, name):
    pr = CASES[name]()
    ia, ja, a, b, tab = oracle.assemble(pr)
    fe, A, bv = gpu.assemble(pr)
    rng = np.random.default_rng(7)
    for _ in range(3):
        x = rng.standard_normal(fe.N)
        y = gpu.Vector(fe.N)
        gpu.prod_matrix_vector(A, gpu.Vector(fe.N, x), y)
        y0 = oracle.amux(ia, ja, a, x)
        assert np.abs(y.get() - y0).max() <= 1e-14 * np.abs(a).max() * np.abs(x).max() * 64


def test_linalg_add_double_path(gpu, oracle):
    """LinAlg_AddDoubleInMatrix / AddDoubleInVector staging: duplicates summed
    in call order into a sorted CSR (add_matrix_double + csr_format)."""
    rng = np.random.default_rng(1)
    n = 40
    A = gpu.Matrix(n)
    dense = np.zeros((n, n))
    for _ in range(600):
        i, j = rng.integers(0, n, 2)
        v = rng.standard_normal()
        A.add_double(i, j, v)
        dense[i, j] += v
    for i in range(n):
        A.add_double(i, i, 10.0)
        dense[i, i] += 10.0
    A.assemble()
    ia, ja, a = A.get_csr()
    M = np.zeros((n, n))
    for i in range(n):
        assert np.all(np.diff(ja[ia[i]:ia[i + 1]]) > 0)
        M[i, ja[ia[i]:ia[i + 1]]] = a[ia[i]:ia[i + 1]]
    assert np.allclose(M, dense, rtol=1e-15, atol=1e-14)
    v = gpu.Vector(n)
    v.add_double(3, 1.5)
    v.add_double(3, 2.0)
    v.assemble()
    assert v.get()[3] == 3.5


def test_spmv_compressed_columns(gpu, oracle, monkeypatch):
    """16-bit column offsets per pair group give the same SpMV bits as the
    32-bit columns; an ordering whose groups span 65536 columns or more (the
    component-blocked elasticity numbering) keeps the 32-bit columns."""
    pr = P.poisson2d(48)
    ia, ja, a, b, tab = oracle.assemble(pr)
    x = np.random.default_rng(3).standard_normal(ia.shape[0] - 1)
    ys = []
    for col32 in ("0", "1"):
        monkeypatch.setenv("GDP_SELL_COL32", col32)
        fe, A, bv = gpu.assemble(pr)
        assert gpu.sell_info(A)["compressed"] == (col32 == "0")
        y = gpu.Vector(fe.N)
        gpu.prod_matrix_vector(A, gpu.Vector(fe.N, x), y)
        ys.append(y.get())
    assert np.array_equal(ys[0], ys[1])
    y0 = oracle.amux(ia, ja, a, x)
    assert np.abs(ys[0] - y0).max() <= 1e-14 * np.abs(a).max() * np.abs(x).max() * 64
    monkeypatch.setenv("GDP_SELL_COL32", "0")
    big = P.elasticity3d(32, nbins=4)          # ux block, then uy, uz: ~36k apart per row
    fe, A, bv = gpu.assemble(big)
    assert not gpu.sell_info(A)["compressed"]


@pytest.mark.parametrize("name", ["poisson2d_lognormal_24", "elasticity3d_4", "curlcurl3d_4",
                                  "poisson2d_az_general"])
def test_device_pre_pass_matches_host(gpu, monkeypatch, name):
    """The PRE pass on the device (numbering, incidence, pattern, assembly
    tables) gives the host pass's DOF tables and pattern, and the same
    assembled system."""
    pr = CASES[name]()
    out = {}
    for host in ("1", "0"):
        monkeypatch.setenv("GDP_HOST_PREP", host)
        fe, A, bv = gpu.assemble(pr)
        ia, ja, a = A.get_csr()
        out[host] = ([fe.dof_table(s) for s in range(len(pr.spaces))], ia, ja, a, bv.get())
    h, d = out["1"], out["0"]
    for s in range(len(pr.spaces)):
        assert np.array_equal(h[0][s], d[0][s])
    for k in range(1, 5):
        assert np.array_equal(h[k], d[k])


@pytest.mark.parametrize("name", list(DIRICHLET_CASES))
def test_solve_with_nonzero_dirichlet(gpu, oracle, name):
    """ILU(0)-GMRES on the lifted systems: solution and iteration count
    against the oracle's."""
    pr = DIRICHLET_CASES[name]()
    print(f"case {name}", flush=True)
    ia, ja, a, b, tab = oracle.assemble(pr)
    fe, A, bv = gpu.assemble(pr)
    s = gpu.Solver(Algorithm=8, Preconditioner=6, Krylov_Size=30)
    x = gpu.Vector(fe.N)
    info = gpu.solve(A, bv, s, x)
    xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30)
    assert info["status"] == 0 and info["rel_res"] <= 1.5e-10
    assert abs(info["its"] - io["its"]) <= max(2, io["its"] // 50), (info["its"], io["its"])
    tol = 1e-8 if name.startswith("curlcurl") else 1e-9
    assert np.abs(x.get() - xo).max() <= tol * np.abs(xo).max()


def fan(m: int, seed: int = 3):
    """A disc of 2D triangles around one centre node: m triangles meet at the
    centre (its matrix row holds m + 1 entries), a second ring of 2 m
    triangles, the outer ring fixed to non-zero values; a mass term and a
    source on part of the disc."""
    rng = np.random.default_rng(seed)
    ang = 2 * np.pi * (np.arange(m) + rng.uniform(-0.2, 0.2, m)) / m
    r1 = 0.5 + rng.uniform(-0.05, 0.05, m)
    xyz = np.zeros((1 + 2 * m, 3))
    xyz[1:m + 1, 0], xyz[1:m + 1, 1] = r1 * np.cos(ang), r1 * np.sin(ang)
    xyz[m + 1:, 0], xyz[m + 1:, 1] = np.cos(ang), np.sin(ang)
    inner = 2 + np.arange(m)           # node numbers (1-based)
    outer = m + 2 + np.arange(m)
    nxt = np.roll(np.arange(m), -1)
    tris = [np.stack([np.ones(m, int), inner, inner[nxt]], 1),
            np.stack([inner, outer, outer[nxt]], 1), np.stack([inner, outer[nxt], inner[nxt]], 1)]
    tri = np.concatenate(tris).astype(np.int32)
    nt = tri.shape[0]
    conn = np.zeros((nt, 4), np.int32)
    conn[:, :3] = tri
    region = np.where(rng.uniform(size=nt) < 0.5, 1, 2).astype(np.int32)
    nu = np.array([[0.0], [1.0], [3.5]])
    ms = np.array([[0.0], [0.2], [0.7]])
    f = np.array([[0.0], [0.0], [-2.0]])
    space = P.Space(P.BF_NODE, 1, [1, 2], outer.astype(np.int32), rng.uniform(-1, 1, m))
    terms = [P.Term(P.T_GRAD_GRAD, 0, 0, [1, 2], 3, nu, dim=2),
             P.Term(P.T_NODE_MASS, 0, 0, [1, 2], 3, ms, dim=2),
             P.Term(P.T_SRC_NODE, 0, -1, [2], 3, f, dim=2)]
    return P.Problem(f"fan_{m}", xyz, np.full(nt, P.TRI, np.int32), region, conn, [space], terms)


@pytest.mark.parametrize("m", [150, 420])
def test_long_rows_bit_exact(gpu, oracle, m):
    """Rows longer than the device PRE pass takes (the centre of a fan of m
    triangles: m + 1 entries): the host PRE pass builds the pattern and the
    contribution lists; assembly bit-exact against the oracle."""
    pr = fan(m)
    print(f"case fan_{m}", flush=True)
    ia, ja, a, b, tab = oracle.assemble(pr)
    assert ia[1] - ia[0] == m + 1
    fe, A, bv = gpu.assemble(pr)
    gia, gja, ga = A.get_csr()
    assert np.array_equal(gia, ia) and np.array_equal(gja, ja)
    assert np.array_equal(fe.dof_table(0), tab[0])
    assert np.array_equal(ga, a)
    assert np.array_equal(bv.get(), b)
