"""The oracle's assembly restatement pinned by closed forms and invariants.

No golden element matrices exist in the reference (its only test is a smoke
run, CMakeLists.txt:1142-1151), so each term is checked against mathematics
that a plausible mistake (dropped quadrature factor, wrong sign, transposed
Jacobian, wrong edge orientation) would violate.
"""
import numpy as np
import pytest

from getdp_amd import problems as P

W3 = 3 * 0.16666666666666      # reference's truncated 3-point triangle weights
W4 = 4 * 0.0416666666667       # reference's truncated 4-point tetrahedron weights


def one_element(xyz, conn, etype, kind, ncoef=1, coef=None, npoints=None, bf=P.BF_NODE,
                edges=None, jac_sur=0, dof=0):
    xyz = np.asarray(xyz, float)
    if xyz.shape[1] == 2:
        xyz = np.c_[xyz, np.zeros(len(xyz))]
    c = np.zeros((1, 4), np.int32)
    c[0, :len(conn)] = conn
    cf = np.asarray(coef if coef is not None else [1.0] * ncoef, float).reshape(1, -1)
    if npoints is None:
        npoints = 3 if etype == P.TRI else 4
    sp = P.Space(bf, 1, [0])
    t = P.Term(kind, 0, dof, [0], npoints, cf, jac_sur=jac_sur)
    return P.Problem("one", xyz, np.array([etype], np.int32), np.array([0], np.int32), c, [sp], [t],
                     edges=edges, edge_nodes=None if edges is None else np.zeros((6, 2), np.int32))


def test_p1_triangle_stiffness_closed_form(oracle):
    pr = one_element([[0, 0], [1, 0], [0, 1]], [1, 2, 3], P.TRI, P.T_GRAD_GRAD)
    K = oracle.element_matrix(pr, 0, 0)
    expect = W3 * np.array([[2, -1, -1], [-1, 1, 0], [-1, 0, 1]]) * 1.0
    assert np.allclose(K, expect, rtol=1e-15, atol=1e-16)
    assert abs(W3 - 0.5) < 1e-13


def test_p1_triangle_stiffness_general_triangle(oracle):
    """K = area * coef * G G^T with G the physical gradients (cotangent formula)."""
    X = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]])
    pr = one_element(X, [1, 2, 3], P.TRI, P.T_GRAD_GRAD, coef=[2.5])
    K = oracle.element_matrix(pr, 0, 0)
    T = np.c_[np.ones(3), X]
    G = np.linalg.inv(T)[1:].T          # rows: gradient of barycentric i
    area = 0.5 * abs(np.linalg.det(T))
    assert np.allclose(K, (W3 / 0.5) * area * 2.5 * G @ G.T, rtol=1e-13, atol=1e-15)


def test_p1_tet_stiffness_and_tensor_form(oracle):
    X = np.array([[0.0, 0.1, 0.0], [1.1, 0.0, 0.2], [0.2, 0.9, 0.1], [0.1, 0.3, 1.2]])
    T = np.c_[np.ones(4), X]
    G = np.linalg.inv(T)[1:].T
    vol = abs(np.linalg.det(T)) / 6
    pr = one_element(X, [1, 2, 3, 4], P.TET, P.T_GRAD_GRAD, coef=[1.7])
    K = oracle.element_matrix(pr, 0, 0)
    assert np.allclose(K, (W4 / (1 / 6)) * vol * 1.7 * G @ G.T, rtol=1e-12)
    # tensor coefficient (asymmetric on purpose): K_ij = g_i . (C g_j)
    C = np.array([[2.0, 0.3, 0.0], [0.1, 1.0, 0.5], [0.0, 0.2, 3.0]])
    pr = one_element(X, [1, 2, 3, 4], P.TET, P.T_GRAD_GRAD, ncoef=9, coef=C.ravel())
    pr.terms[0].dof = 0
    K = oracle.element_matrix(pr, 0, 0)
    # symmetric term (same quantity): the reference evaluates j <= i and
    # mirrors, i.e. K_ij = g_max(i,j) . C g_min(i,j)
    ref = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            a, b = max(i, j), min(i, j)
            ref[i, j] = G[a] @ (C @ G[b])
    assert np.allclose(K, (W4 / (1 / 6)) * vol * ref, rtol=1e-12)


def test_elasticity_rigid_modes_in_kernel(oracle):
    """Without constraints the elasticity operator annihilates the six rigid
    body motions (3 translations, 3 rotations)."""
    pr = P.elasticity3d(2, nbins=4)
    for s in pr.spaces:
        s.fixed_entity = np.zeros(0, np.int32)
        s.fixed_value = np.zeros(0)
    ia, ja, a, b, tab = oracle.assemble(pr)
    N = ia.shape[0] - 1
    xyz = pr.xyz
    modes = []
    for d in range(3):
        u = np.zeros((pr.nnodes, 3)); u[:, d] = 1.0; modes.append(u)
    for (p, q) in ((0, 1), (1, 2), (2, 0)):
        u = np.zeros((pr.nnodes, 3)); u[:, p] = -xyz[:, q]; u[:, q] = xyz[:, p]; modes.append(u)
    for u in modes:
        v = np.zeros(N)
        for s in range(3):
            d = tab[s, 1:pr.nnodes + 1]
            v[d[d > 0] - 1] = u[d > 0, s]
        r = oracle.amux(ia, ja, a, v)
        assert np.abs(r).max() < 1e-12 * np.abs(a).max()
    # block DOF numbering: all ux, then all uy, then all uz (F7 contract)
    assert tab[0, 1:].max() < tab[1, 1:][tab[1, 1:] > 0].min()
    assert tab[1, 1:].max() < tab[2, 1:][tab[2, 1:] > 0].min()


def test_poisson_neumann_row_sums_and_source(oracle):
    pr = P.poisson2d(8)
    pr.spaces[0].fixed_entity = np.zeros(0, np.int32)
    pr.spaces[0].fixed_value = np.zeros(0)
    ia, ja, a, b, tab = oracle.assemble(pr)
    N = ia.shape[0] - 1
    assert N == pr.nnodes
    rows = oracle.amux(ia, ja, a, np.ones(N))
    assert np.abs(rows).max() < 1e-14
    # symmetric matrix with the full element-adjacency pattern, zeros stored
    A = np.zeros((N, N))
    for i in range(N):
        A[i, ja[ia[i]:ia[i + 1]]] = a[ia[i]:ia[i + 1]]
    assert np.allclose(A, A.T, rtol=0, atol=0)
    assert (a == 0).sum() > 0  # hypotenuse couplings of the right triangles are exact zeros
    # sum of b = js * area of the source region (partition of unity)
    ntri = pr.meta["ntri"]
    src = pr.region[:ntri] > pr.meta["ncls"]
    tri = pr.conn[:ntri, :3] - 1
    X = pr.xyz
    ar = 0.5 * np.abs(np.cross(X[tri[:, 1]] - X[tri[:, 0]], X[tri[:, 2]] - X[tri[:, 0]])[:, 2])
    assert abs(b.sum() - W3 / 0.5 * ar[src].sum()) < 1e-12


def test_curl_curl_kills_gradients_and_edge_mass(oracle):
    """curl grad = 0: the curl-curl matrix maps the edge DOFs of any discrete
    gradient (phi(hi) - phi(lo) on each edge lo->hi) to zero; the edge mass
    reproduces |t|^2 |Omega| for a constant field t."""
    pr = P.curlcurl3d(2)
    pr.spaces[0].fixed_entity = np.zeros(0, np.int32)
    pr.spaces[0].fixed_value = np.zeros(0)
    cc, ms, src = pr.terms
    rng = np.random.default_rng(0)
    phi = rng.standard_normal(pr.nnodes)
    en = pr.edge_nodes
    pr.terms = [cc]
    ia, ja, a, b, tab = oracle.assemble(pr)
    N = ia.shape[0] - 1
    d = tab[0, 1:en.shape[0] + 1]
    g = np.zeros(N)
    g[d - 1] = phi[en[:, 1] - 1] - phi[en[:, 0] - 1]
    assert np.abs(oracle.amux(ia, ja, a, g)).max() < 1e-11 * np.abs(a).max()
    # mass term alone, unit coefficient
    ms.coef = np.where(ms.coef != 0, 1.0, 0.0)
    pr.terms = [ms]
    ia, ja, a, b, tab = oracle.assemble(pr)
    t = np.array([0.3, -1.2, 0.7])
    X = pr.xyz
    dofs = np.zeros(N)
    dofs[d - 1] = (X[en[:, 1] - 1] - X[en[:, 0] - 1]) @ t
    energy = dofs @ oracle.amux(ia, ja, a, dofs)
    assert abs(energy - (t @ t) * 1.0 * (W4 / (1 / 6))) < 1e-10


def test_edge_source_integrates_field(oracle):
    """[ f, {a} ] with constant f: b . dofs(t) = -(f . t) |Omega|."""
    pr = P.curlcurl3d(2)
    pr.spaces[0].fixed_entity = np.zeros(0, np.int32)
    pr.spaces[0].fixed_value = np.zeros(0)
    f = np.array([0.2, -0.5, 1.5])
    src = pr.terms[2]
    src.coef = np.tile(f, (src.coef.shape[0], 1))
    src.regions = [1, 2, 3, 4]
    pr.terms = [pr.terms[1], src]
    ia, ja, a, b, tab = oracle.assemble(pr)
    en = pr.edge_nodes
    d = tab[0, 1:en.shape[0] + 1]
    t = np.array([1.0, 2.0, -0.5])
    dofs = np.zeros(ia.shape[0] - 1)
    dofs[d - 1] = (pr.xyz[en[:, 1] - 1] - pr.xyz[en[:, 0] - 1]) @ t
    assert abs(b @ dofs + (f @ t) * (W4 / (1 / 6))) < 1e-12


def test_traction_surface_term(oracle):
    """Unit traction on x=1: the ux equations receive -sum = -area(face)."""
    pr = P.elasticity3d(3, nbins=2)
    ia, ja, a, b, tab = oracle.assemble(pr)
    assert abs(b.sum() + W3 / 0.5 * 1.0) < 1e-12
    nx = (tab[0] > 0).sum()
    assert np.all(b[nx:] == 0.0)


def test_edge_numbering_hand_computed(oracle):
    """First-encounter edge numbering with the reference's local edge table
    of the tetrahedron (GeoEntity.h:214-215: (0,1) (0,2) (0,3) (1,2) (1,3)
    (2,3)) and the sign rule (+ when the local edge runs from the lower to the
    higher node number), worked out by hand for two tetrahedra sharing a face,
    visited in both orders."""
    from getdp_amd.problems import TET
    conn = np.array([[1, 2, 3, 4], [5, 3, 2, 4]], np.int32)
    etype = np.array([TET, TET], np.int32)
    e, en = oracle.number_edges(etype, conn, np.array([0, 1], np.int32))
    assert e.tolist() == [[1, 2, 3, 4, 5, 6], [-7, -8, -9, -4, 6, 5]]
    assert [tuple(sorted(p)) for p in en.tolist()] == [(1, 2), (1, 3), (1, 4), (2, 3), (2, 4), (3, 4),
                                                      (3, 5), (2, 5), (4, 5)]
    e, en = oracle.number_edges(etype, conn, np.array([1, 0], np.int32))
    assert e.tolist() == [[7, 8, 9, 4, 6, 5], [-1, -2, -3, -4, 5, 6]]
    assert [tuple(sorted(p)) for p in en.tolist()] == [(3, 5), (2, 5), (4, 5), (2, 3), (3, 4), (2, 4),
                                                      (1, 2), (1, 3), (1, 4)]


def test_dof_numbering_hand_computed(oracle):
    """DOF numbers follow the key order (NumType, Entity) of fcmp_Dof
    (DofData.cpp:36-45), fixed DOFs get none: on one cube (8 nodes, x = 0
    clamped), the 4 free nodes 2, 4, 6, 8 are numbered 1-4 for u_x, 5-8 for
    u_y and 9-12 for u_z; clamped nodes carry -(index into the fixed list + 1)."""
    import getdp_amd.problems as P
    pr = P.elasticity3d(1, nbins=4)
    N, tab = oracle.number_dofs(pr)
    assert N == 12
    free = [n for n in range(1, 9) if pr.xyz[n - 1, 0] > 0.5]
    assert free == [2, 4, 6, 8]
    for s in range(3):
        for r, node in enumerate(free):
            assert tab[s, node] == 4 * s + r + 1
        for node in range(1, 9):
            if node not in free:
                f = list(pr.spaces[s].fixed_entity).index(node)
                assert tab[s, node] == -(f + 1)


def _rot(g):
    """curl (phi e_z) in the plane: (d phi/dy, -d phi/dx, 0)."""
    return np.array([g[1], -g[0], 0.0])


def test_perpendicular_edge_closed_forms(oracle):
    """BF_PerpendicularEdge (a_z): curl-curl = area * (R g_i) . C (R g_j) with
    R the -90 degree turn of the node gradients; mass = area (1 + d_ij) / 12;
    source = f_z area / 3 (only the z component enters, Cal_Product3).  The
    anisotropic tensor tells a turned gradient from a swapped one."""
    X = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]])
    T = np.c_[np.ones(3), X]
    G = np.linalg.inv(T)[1:].T
    G3 = np.c_[G, np.zeros(3)]
    area = 0.5 * abs(np.linalg.det(T))
    s = W3 / 0.5
    pr = one_element(X, [1, 2, 3], P.TRI, P.T_CURL_CURL, coef=[2.5], bf=P.BF_PERP_EDGE)
    K = oracle.element_matrix(pr, 0, 0)
    assert np.allclose(K, s * area * 2.5 * G @ G.T, rtol=1e-13, atol=1e-15)
    C = np.array([[2.0, 0.7, 0.0], [-0.4, 1.0, 0.0], [0.0, 0.0, 3.0]])
    pr = one_element(X, [1, 2, 3], P.TRI, P.T_CURL_CURL, ncoef=9, coef=C.ravel(), bf=P.BF_PERP_EDGE)
    K = oracle.element_matrix(pr, 0, 0)
    ref = np.array([[_rot(G3[max(i, j)]) @ C @ _rot(G3[min(i, j)]) for j in range(3)] for i in range(3)])
    assert np.allclose(K, s * area * ref, rtol=1e-13, atol=1e-15)
    swapped = np.array([[G3[max(i, j)][[1, 0, 2]] @ C @ G3[min(i, j)][[1, 0, 2]] for j in range(3)]
                        for i in range(3)])
    assert not np.allclose(K, s * area * swapped, rtol=1e-6)
    pr = one_element(X, [1, 2, 3], P.TRI, P.T_EDGE_MASS, coef=[0.3], bf=P.BF_PERP_EDGE)
    M = oracle.element_matrix(pr, 0, 0)
    assert np.allclose(M, s * area * 0.3 * (np.ones((3, 3)) + np.eye(3)) / 12, rtol=1e-12)
    pr = one_element(X, [1, 2, 3], P.TRI, P.T_SRC_EDGE, ncoef=3, coef=[5.0, -7.0, -1.5],
                     bf=P.BF_PERP_EDGE, dof=-1)
    f = oracle.element_matrix(pr, 0, 0).ravel()
    assert np.allclose(f, s * area * -1.5 / 3, rtol=1e-12)


def test_perpendicular_edge_is_the_nodal_problem(oracle):
    """The a_z formulation of the 2D problem assembles the nodal Poisson
    system: same DOF numbering, pattern and right-hand side, matrix entries
    equal up to rounding (J/detJ against J^-1)."""
    rng = np.random.default_rng(5)
    prs = [P.poisson2d(10, seed=4, field_kind="lognormal", sigma=1.0, nbins=8, basis=b)
           for b in ("node", "perp")]
    inner = np.all((prs[0].xyz[:, :2] > 0) & (prs[0].xyz[:, :2] < 1), axis=1)
    jit = rng.uniform(-0.03, 0.03, size=(int(inner.sum()), 2))
    for pr in prs:
        pr.xyz[inner, :2] += jit
    (ia, ja, a, b, tab), (ia2, ja2, a2, b2, tab2) = (oracle.assemble(pr) for pr in prs)
    assert np.array_equal(ia, ia2) and np.array_equal(ja, ja2) and np.array_equal(tab, tab2)
    assert np.array_equal(b, b2)
    assert np.abs(a - a2).max() <= 1e-13 * np.abs(a).max()
    assert not np.array_equal(a, a2)      # the two change-of-coordinate paths round differently


# ---------------------------------------------------------------------------
# Fixed DOFs with non-zero values: the first constraint holding an entity wins
# (Get_ConstraintOfElement.cpp:51-54, DofData.cpp:1137-1151), and the
# unknown x fixed branch of Dof_AssembleInMat lifts b -= value * Ek
# (DofData.cpp:1489-1500).


def test_first_constraint_wins_hand_worked(oracle):
    """3x3 nodes, 8 right triangles, unit reluctivity, no source: node 5 (the
    centre) is the only unknown.  Its row of the P1 stiffness is the 5-point
    Laplacian scaled by W3/0.5 (the hypotenuse couplings to nodes 1 and 9 are
    exact zeros), so b_5 = (W3/0.5) * (v2 + v4 + v6 + v8) with the FIRST value
    listed for each node."""
    pr = P.poisson2d(2, nu_vals=(1.0, 1.0), js=0.0)
    ent = [2, 4, 6, 8, 1, 3, 7, 9, 2, 6, 4]
    val = [1.0, 2.0, 3.0, 4.0, 7.0, 7.0, 7.0, 7.0, 100.0, -50.0, 9.0]
    pr.spaces[0].fixed_entity = np.array(ent, np.int32)
    pr.spaces[0].fixed_value = np.array(val)
    ia, ja, a, b, tab = oracle.assemble(pr)
    assert b.shape == (1,) and tab[0][5] == 1
    # each boundary node keeps the index of its first occurrence
    for node, first in {2: 0, 4: 1, 6: 2, 8: 3, 1: 4, 3: 5, 7: 6, 9: 7}.items():
        assert tab[0][node] == -(first + 1)
    s = W3 / 0.5
    assert a[0] == pytest.approx(4 * s, rel=1e-15)
    assert b[0] == pytest.approx(s * (1 + 2 + 3 + 4), rel=1e-14)
    # a last-wins rule would give s * (100 + 9 - 50 + 4)
    assert abs(b[0] - s * 63) > 1.0


def _solve_with_fixed(pr, oracle):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    ia, ja, a, b, tab = oracle.assemble(pr)
    N = b.shape[0]
    A = sp.csr_matrix((a, ja, ia), shape=(N, N))
    x = spla.spsolve(A.tocsc(), b)
    return x, tab


def test_dirichlet_patch_test_poisson(oracle):
    """An affine field is the exact discrete solution of the Laplace problem
    with its own boundary values (P1 reproduces linear functions), so every
    node must carry it to rounding; the lifting of non-zero fixed values is
    the only thing that drives the solution."""
    n = 9
    pr = P.poisson2d(n, nu_vals=(2.5, 2.5), js=0.0)
    sp0 = pr.spaces[0]
    u = lambda X: 1.0 + 2.0 * X[:, 0] - 3.0 * X[:, 1]
    sp0.fixed_value = u(pr.xyz[sp0.fixed_entity - 1])
    x, tab = _solve_with_fixed(pr, oracle)
    t = tab[0]
    inner = np.flatnonzero(t > 0)
    got = x[t[inner] - 1]
    assert np.abs(got - u(pr.xyz[inner - 1])).max() <= 1e-12
    assert inner.size == (n - 1) ** 2


def test_dirichlet_patch_test_elasticity(oracle):
    """Constant strain: u = c + G x with uniform material is in equilibrium
    with zero body force, so prescribing it on the whole boundary of the cube
    must reproduce it at every interior node (all three block-numbered
    components)."""
    n = 3
    pr = P.elasticity3d(n, sigma=0.0, traction=0.0)
    X = pr.xyz
    bnd = (np.flatnonzero(np.any((X == 0.0) | (X == 1.0), axis=1)) + 1).astype(np.int32)
    G = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [-0.1, 0.2, 0.25]])
    c = np.array([1.0, -2.0, 0.5])
    U = c + X @ G.T
    for s in range(3):
        pr.spaces[s].fixed_entity = bnd
        pr.spaces[s].fixed_value = U[bnd - 1, s]
    x, tab = _solve_with_fixed(pr, oracle)
    for s in range(3):
        inner = np.flatnonzero(tab[s] > 0)
        assert inner.size == (n - 1) ** 3
        assert np.abs(x[tab[s][inner] - 1] - U[inner - 1, s]).max() <= 1e-12


@pytest.mark.parametrize("family", ["poisson", "elasticity", "curlcurl", "az"])
def test_dirichlet_lifting_equals_unconstrained_system(oracle, family):
    """The same problem assembled without constraints gives the full
    operator K and load f over every entity; with fixed entities F (values
    v, the first listed value of each) the constrained system must be
    A = K[U, U] and b = f[U] - K[U, F] v.  Checks the lifting of non-zero
    values through a different code path (unknown x unknown only) for every
    family, with entities listed twice under different values."""
    import copy
    from dirichlet import first_values, with_dirichlet
    make = {"poisson": lambda: P.poisson2d(6, field_kind="lognormal", nbins=4),
            "elasticity": lambda: P.elasticity3d(2, nbins=4),
            "curlcurl": lambda: P.curlcurl3d(3),
            "az": lambda: P.poisson2d(6, basis="perp")}[family]
    pr = with_dirichlet(make(), seed=5, ndup=8)
    free = copy.deepcopy(pr)
    for sp_ in free.spaces:
        sp_.fixed_entity = np.zeros(0, np.int32)
        sp_.fixed_value = np.zeros(0)
    ia, ja, a, b, tab = oracle.assemble(pr)
    fia, fja, fa, fb, ftab = oracle.assemble(free)
    Nf = fb.shape[0]
    K = np.zeros((Nf, Nf))
    for r in range(Nf):
        K[r, fja[fia[r]:fia[r + 1]]] += fa[fia[r]:fia[r + 1]]
    U, F, vF = [], [], []
    for si, sp_ in enumerate(pr.spaces):
        fv = first_values(sp_)
        for ent in np.flatnonzero(ftab[si] > 0):
            if tab[si][ent] > 0:
                U.append((tab[si][ent] - 1, ftab[si][ent] - 1))
            else:
                assert tab[si][ent] < 0 and ent in fv
                # the constrained table points at the first occurrence
                assert sp_.fixed_value[-tab[si][ent] - 1] == fv[ent]
                F.append(ftab[si][ent] - 1)
                vF.append(fv[ent])
    U.sort()
    urow = np.array([u[1] for u in U])
    assert len(U) == b.shape[0] and len(F) > 0
    N = b.shape[0]
    A = np.zeros((N, N))
    for r in range(N):
        A[r, ja[ia[r]:ia[r + 1]]] += a[ia[r]:ia[r + 1]]
    assert np.array_equal(A, K[np.ix_(urow, urow)])
    expect = fb[urow] - K[np.ix_(urow, F)] @ np.array(vF)
    scale = np.abs(K[np.ix_(urow, F)]).max() * np.abs(vF).max() + np.abs(fb).max()
    assert np.abs(b - expect).max() <= 1e-13 * scale
