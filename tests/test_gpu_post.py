"""Device post-processing (gdp_fe_post) against the oracle's restatement
(orc_post, itself pinned by exact reproduction of fields in the
interpolation spaces and by the energy identity, tests/test_oracle_post.py).

Bar: every per-element value bit-identical (same operations in the same
order); the integral's total, a device reduction in a different order from
the reference's element-order sum, within 1e-13 relative."""
import numpy as np
import pytest

from getdp_amd import problems as P

pytestmark = pytest.mark.gpu

CASES = {
    "poisson2d_16": lambda: P.poisson2d(16),
    "poisson2d_16_perp": lambda: P.poisson2d(16, basis="perp"),
    "elasticity3d_4": lambda: P.elasticity3d(4, nbins=16),
    "curlcurl3d_4": lambda: P.curlcurl3d(4),
}


def _regions(pr):
    """Regions of the problem's volume elements (triangles in 2D, tetrahedra in 3D)."""
    vol = P.TET if (pr.etype == P.TET).any() else P.TRI
    return sorted(set(int(r) for r, t in zip(pr.region, pr.etype) if t == vol))


@pytest.mark.parametrize("name", list(CASES))
def test_post_matches_oracle(gpu, oracle, name):
    print(f"case {name}", flush=True)
    pr = CASES[name]()
    rng = np.random.default_rng(4)
    for sp in pr.spaces:  # non-zero constraint values exercise the fixed DOFs
        sp.fixed_value = rng.standard_normal(sp.fixed_entity.shape[0])
    fe = gpu.FESystem(pr)
    x = rng.standard_normal(fe.N)
    xv = gpu.Vector(fe.N, x)
    regs = _regions(pr)
    nreg = int(pr.region.max()) + 1
    scal = np.linspace(0.5, 2.0, nreg)[:, None]
    tens = rng.standard_normal((nreg, 9))
    for space in range(len(pr.spaces)):
        for kind, coef in ((gpu.PQ_VALUE, None), (gpu.PQ_D, None), (gpu.PQ_COEF_D, scal),
                           (gpu.PQ_COEF_D, tens)):
            got = fe.post(xv, kind, regs, space=space, coef=coef)
            want = oracle.post(pr, x, kind, regs, space=space, coef=coef)
            assert got["nsel"] == want["nsel"]
            assert np.array_equal(got["values"].view(np.uint64), want["values"].view(np.uint64)), \
                (name, space, kind, np.abs(got["values"] - want["values"]).max())
        for npoints in ((1, 3) if pr.etype[0] == P.TRI else (1, 4)):
            got = fe.post(xv, gpu.PQ_INT_COEF_D2, regs, space=space, coef=scal, npoints=npoints)
            want = oracle.post(pr, x, 4, regs, space=space, coef=scal, npoints=npoints)
            assert np.array_equal(got["values"].view(np.uint64), want["values"].view(np.uint64))
            assert abs(got["total"] - want["total"]) <= 1e-13 * abs(want["total"])


def test_post_energy_of_the_solution(gpu, oracle):
    """Print[energy[Dom], OnGlobal] on the device solution equals x^T A x."""
    pr = P.poisson2d(32)
    fe, A, b = gpu.assemble(pr)
    s = gpu.Solver(Algorithm=8, Preconditioner=6, Krylov_Size=30)
    x = gpu.Vector(fe.N)
    gpu.solve(A, b, s, x)
    t = pr.terms[0]
    r = fe.post(x, gpu.PQ_INT_COEF_D2, t.regions, coef=t.coef, npoints=t.npoints)
    xs = x.get()
    y = gpu.Vector(fe.N)
    gpu.prod_matrix_vector(A, x, y)
    xAx = float(xs @ y.get())
    assert abs(r["total"] - xAx) <= 1e-12 * abs(xAx)


def test_post_rejects_bad_requests(gpu):
    pr = P.poisson2d(8)
    fe = gpu.FESystem(pr)
    x = gpu.Vector(fe.N)
    from getdp_amd._lib import GdpError
    with pytest.raises(GdpError):
        fe.post(x, 9, [1])                          # unknown quantity
    with pytest.raises(GdpError):
        fe.post(x, gpu.PQ_COEF_D, [1], coef=None)   # coefficient missing
    with pytest.raises(GdpError):
        fe.post(x, gpu.PQ_D, [0])                   # boundary lines: not a volume element
