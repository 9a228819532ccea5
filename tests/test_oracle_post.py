"""The oracle's post-processing restatement (orc_post: Pos_LocalOrIntegralQuantity
+ Pos_FemInterpolation) pinned by what the mathematics fixes:

* P1 nodal, Whitney edge and 2D perpendicular-edge interpolants reproduce
  fields of their own spaces exactly: a linear u (nodal, a_z), a + b x x
  (lowest-order Nedelec).  Values at the vertices and the constant
  gradient / curl must come out to rounding.
* The energy integral sum_e Integral coef |grad u|^2 equals x^T A x for the
  assembled grad-grad matrix (Dirichlet values zero), a different
  computation of the same number.
"""
import numpy as np
import pytest

from getdp_amd import problems as P

VALUE, D, COEF_D, INT = 1, 2, 3, 4


def _dof_vector(oracle, pr, f_entity):
    """x for the unknowns and constraint values of space 0 from per-entity values."""
    sp = pr.spaces[0]
    sp.fixed_value = np.asarray(f_entity[sp.fixed_entity], np.float64)
    N, tab = oracle.number_dofs(pr)
    x = np.zeros(N)
    for ent in range(1, tab.shape[1]):
        c = tab[0, ent]
        if c > 0:
            x[c - 1] = f_entity[ent]
    return x


def _vol_regions(pr):
    return sorted(set(int(r) for r, t in zip(pr.region, pr.etype) if t in (P.TRI, P.TET)) - {0})


def test_nodal_linear_field_is_reproduced(oracle):
    pr = P.poisson2d(8)
    a, c0 = np.array([0.7, -1.3, 0.0]), 0.25
    f = np.zeros(pr.nnodes + 1)
    f[1:] = pr.xyz @ a + c0
    x = _dof_vector(oracle, pr, f)
    regs = _vol_regions(pr)
    r = oracle.post(pr, x, D, regs)
    assert r["nsel"] == int((pr.etype == P.TRI).sum())
    g = r["values"][:, :3, :]
    assert np.allclose(g, a, rtol=0, atol=1e-12)
    assert np.all(r["values"][:, 3, :] == 0)
    v = oracle.post(pr, x, VALUE, regs)["values"]
    tris = pr.conn[pr.etype == P.TRI][:, :3]
    assert np.allclose(v[:, :3, 0], f[tris], rtol=0, atol=1e-14)
    assert np.all(v[:, :, 1:] == 0)


def test_perpendicular_edge_linear_az(oracle):
    pr = P.poisson2d(8, basis="perp")
    a, c0 = np.array([0.4, 0.9, 0.0]), -0.1
    f = np.zeros(pr.nnodes + 1)
    f[1:] = pr.xyz @ a + c0
    x = _dof_vector(oracle, pr, f)
    regs = _vol_regions(pr)
    b = oracle.post(pr, x, D, regs)["values"][:, :3, :]
    # b = curl (a_z e_z) = (d a_z / dy, -d a_z / dx, 0)
    assert np.allclose(b, [a[1], -a[0], 0.0], rtol=0, atol=1e-12)
    v = oracle.post(pr, x, VALUE, regs)["values"]
    tris = pr.conn[pr.etype == P.TRI][:, :3]
    assert np.allclose(v[:, :3, 2], f[tris], rtol=0, atol=1e-14)
    assert np.all(v[:, :, :2] == 0)


def test_whitney_interpolant_of_nedelec_field(oracle):
    pr = P.curlcurl3d(3)
    c0, B = np.array([0.3, -0.2, 0.5]), np.array([1.1, -0.6, 0.8])

    def field(p):
        return c0 + 0.5 * np.cross(B, p)

    lo, hi = pr.edge_nodes[:, 0] - 1, pr.edge_nodes[:, 1] - 1
    mid = 0.5 * (pr.xyz[lo] + pr.xyz[hi])
    f = np.zeros(pr.edge_nodes.shape[0] + 1)
    f[1:] = np.einsum("ij,ij->i", field(mid), pr.xyz[hi] - pr.xyz[lo])  # exact line integral
    x = _dof_vector(oracle, pr, f)
    regs = _vol_regions(pr)
    r = oracle.post(pr, x, D, regs)
    assert np.allclose(r["values"], B, rtol=0, atol=1e-12)
    v = oracle.post(pr, x, VALUE, regs)["values"]
    tets = pr.conn[pr.etype == P.TET]
    assert np.allclose(v, field(pr.xyz[tets - 1]), rtol=0, atol=1e-12)


def test_coef_times_gradient_scalar_and_tensor(oracle):
    pr = P.poisson2d(8)
    x = np.random.default_rng(1).standard_normal(oracle.number_dofs(pr)[0])
    regs = _vol_regions(pr)
    g = oracle.post(pr, x, D, regs)["values"]
    nreg = int(pr.region.max()) + 1
    s = np.linspace(0.5, 2.0, nreg)[:, None]
    gs = oracle.post(pr, x, COEF_D, regs, coef=s)["values"]
    rid = pr.region[pr.etype == P.TRI]
    assert np.array_equal(gs, s[rid][:, :, None] * g)
    T = np.random.default_rng(2).standard_normal((nreg, 9))
    gt = oracle.post(pr, x, COEF_D, regs, coef=T)["values"]
    want = np.einsum("eij,evj->evi", T[rid].reshape(-1, 3, 3), g)
    assert np.allclose(gt, want, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("npoints", [1, 3])
def test_energy_integral_equals_xT_A_x(oracle, npoints):
    pr = P.poisson2d(8, npoints=npoints)
    ia, ja, a, b, tab = oracle.assemble(pr)
    x = np.random.default_rng(3).standard_normal(ia.shape[0] - 1)
    nu = pr.terms[0].coef
    r = oracle.post(pr, x, INT, pr.terms[0].regions, coef=nu, npoints=npoints)
    xAx = float(x @ oracle.amux(ia, ja, a, x))
    assert abs(r["total"] - xAx) <= 1e-12 * abs(xAx)
    assert np.all(r["values"] >= 0)
    assert r["total"] == pytest.approx(float(np.sum(r["values"])), rel=1e-14)


def test_elements_outside_the_support_give_zero(oracle):
    pr = P.curlcurl3d(2)
    sp = pr.spaces[0]
    sp.support = [1, 2]  # regions 3 and 4 now outside the space's support
    x = np.ones(oracle.number_dofs(pr)[0])
    r = oracle.post(pr, x, D, [1, 2, 3, 4])
    rid = pr.region[pr.etype == P.TET]
    assert np.all(r["values"][np.isin(rid, [3, 4])] == 0)
