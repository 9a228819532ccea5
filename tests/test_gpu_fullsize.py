"""Device path against the oracle at the full sizes of the five workloads.

* C0: 2D Poisson 256^2 (131,072 triangles), ILU(0)-GMRES(30) in RCMK order:
  bit-exact assembly, the oracle's iteration count, solution within 1e-9.
* C1: 3D elasticity 64^3 (1,572,864 tetrahedra, 811,200 DOFs, nnz 35.6 M):
  bit-exact assembly (16-bit column offsets, 35.6 M-entry indexing), the
  60-iteration CG/Jacobi iterate against the oracle's, full CG/Jacobi and
  CG/ILU(0) to 1e-10.
* C2: curl-curl 48^3 (663,552 tetrahedra, 753,552 edge DOFs): bit-exact
  assembly, CG/Jacobi to 1e-10 against the reference's own FP64 Sparskit
  solution (tests/golden/c2_curlcurl48_cg.npz), and the error of the nested
  mixed-precision solve against it, stated.
* C3: 512 uneven instances (sides 32..192) in one graph: every instance's
  true residual, eight instances against the oracle's solve.
"""
import numpy as np
import pytest

from getdp_amd import problems as P

pytestmark = pytest.mark.gpu

GOLDEN = __import__("pathlib").Path(__file__).with_name("golden")


def _assembly_bit_exact(gpu, oracle, pr):
    ia, ja, a, b, tab = oracle.assemble(pr)
    fe, A, bv = gpu.assemble(pr)
    gia, gja, ga = A.get_csr()
    assert np.array_equal(gia, ia) and np.array_equal(gja, ja)
    for s in range(len(pr.spaces)):
        assert np.array_equal(fe.dof_table(s), tab[s])
    assert np.array_equal(ga, a)
    assert np.array_equal(bv.get(), b)
    return ia, ja, a, b, fe, A, bv


def test_c0_poisson_256(gpu, oracle):
    print("case C0 poisson2d_256", flush=True)
    pr = P.poisson2d(256)
    ia, ja, a, b, fe, A, bv = _assembly_bit_exact(gpu, oracle, pr)
    assert fe.N == 65025 and ja.shape[0] == 453137
    s = gpu.Solver(Algorithm=8, Preconditioner=6, Krylov_Size=30, Renumbering_Technique=1,
                   Nb_Iter_Max=100000, Stopping_Test=1e-10)
    x = gpu.Vector(fe.N)
    info = gpu.solve(A, bv, s, x)
    xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30, maxits=100000)
    print(f"C0: device its {info['its']} oracle its {io['its']} rel_res {info['rel_res']:.3e}", flush=True)
    assert info["status"] == 0 and io["status"] == 0
    assert info["rel_res"] <= 1.5e-10
    assert abs(info["its"] - io["its"]) <= max(2, io["its"] // 50), (info["its"], io["its"])
    assert np.abs(x.get() - xo).max() <= 1e-9 * np.abs(xo).max()


@pytest.fixture(scope="module")
def c1(gpu, oracle):
    print("case C1 elasticity3d_64 assembly", flush=True)
    pr = P.elasticity3d(64)
    return _assembly_bit_exact(gpu, oracle, pr)


def test_c1_elasticity_64_assembly_and_cg_iterate(gpu, oracle, c1):
    ia, ja, a, b, fe, A, bv = c1
    assert fe.N == 811200 and ja.shape[0] == 35607726
    # 60 CG iterations on the scaled system: the same iterate up to the
    # rounding of the dot products (the device sums them in another order)
    s = gpu.Solver(Algorithm=1, Preconditioner=0, Scaling=1, Renumbering_Technique=1, Nb_Iter_Max=60,
                   Stopping_Test=1e-10)
    x = gpu.Vector(fe.N)
    info = gpu.solve(A, bv, s, x)
    xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=1, preconditioner=0, scaling=1, maxits=60)
    assert info["its"] == io["its"] == 60 and info["status"] == io["status"] == -1
    err = np.abs(x.get() - xo).max() / np.abs(xo).max()
    # how far rounding alone moves the oracle's own 60th iterate: the same
    # algorithm with its dot products summed in 512 blocks, as the device's
    # reductions are (a different rounding order, nothing else)
    oracle.set_dot_blocks(512)
    try:
        xp, ip, _ = oracle.solve(ia, ja, a, b, algorithm=1, preconditioner=0, scaling=1, maxits=60)
    finally:
        oracle.set_dot_blocks(0)
    sens = np.abs(xp - xo).max() / np.abs(xo).max()
    print(f"C1: 60-iteration CG iterate differs from the oracle's by {err:.2e}; blocked dot products "
          f"move the oracle's own by {sens:.2e}", flush=True)
    assert sens > 0 and err <= 10 * sens


def test_c1_elasticity_64_cg_jacobi_and_ilu0_to_1e10(gpu, oracle, c1):
    ia, ja, a, b, fe, A, bv = c1
    d = np.array(a[np.flatnonzero(ja == np.repeat(np.arange(fe.N), np.diff(ia)))])
    w = 1.0 / np.sqrt(np.abs(d))
    for prec in (0, 6):
        s = gpu.Solver(Algorithm=1, Preconditioner=prec, Scaling=1, Renumbering_Technique=1,
                       Nb_Iter_Max=20000, Stopping_Test=1e-10)
        x = gpu.Vector(fe.N)
        info = gpu.solve(A, bv, s, x)
        print(f"C1 CG prec {prec}: its {info['its']} rel_res {info['rel_res']:.3e} "
              f"{info['t_solve_ms']:.0f} ms", flush=True)
        assert info["status"] == 0
        # the reference stops on the scaled system's residual
        xs = x.get()
        r = oracle.amux(ia, ja, a, xs) - b
        assert np.linalg.norm(w * r) <= 1.5e-10 * np.linalg.norm(w * b)
        assert info["rel_res"] <= 2e-10


def test_c2_curlcurl_48(gpu, oracle):
    print("case C2 curlcurl3d_48", flush=True)
    ref = np.load(GOLDEN / "c2_curlcurl48_cg.npz")
    pr = P.curlcurl3d(48)
    ia, ja, a, b, fe, A, bv = _assembly_bit_exact(gpu, oracle, pr)
    assert fe.N == int(ref["n"]) == 753552
    st = int(ref["stride"])
    xr = ref["x_sampled"]
    scale = float(ref["norm_inf"])
    s = gpu.Solver(Algorithm=1, Preconditioner=0, Scaling=1, Renumbering_Technique=1,
                   Nb_Iter_Max=200000, Stopping_Test=1e-10)
    x = gpu.Vector(fe.N)
    info = gpu.solve(A, bv, s, x)
    xs = x.get()
    err_cg = np.abs(xs[::st] - xr).max() / scale
    print(f"C2 CG/Jacobi: its {info['its']} (reference {int(ref['its'])}) rel_res {info['rel_res']:.3e}; "
          f"difference from the reference's FP64 solution {err_cg:.2e}", flush=True)
    # the stop test applies to the scaled residual; the reference's own
    # unscaled residual at exit is 9.4e-10
    assert info["status"] == 0 and info["rel_res"] <= 1e-9
    print(f"C2 norms: device {np.linalg.norm(xs):.6e} reference {float(ref['norm2']):.6e}", flush=True)
    # cond(A) ~ 1e9 with the 1e-6 mass: the two CG runs (different dot-product
    # orders, 51,918 and 48,909 iterations) agree to 2.2e-4 of max |x|
    assert err_cg <= 1e-3
    assert abs(np.linalg.norm(xs) - float(ref["norm2"])) <= 1e-4 * float(ref["norm2"])
    # nested mixed precision (FP64 FGMRES around FP32 GMRES(20) with FP32
    # ILU(0)): stagnates near 6e-7 relative residual on this operator, as the
    # FP64 ILU(0)-GMRES does; its error against the reference's solution is
    # stated, not bounded below the stagnation level
    # (multicolour ILU(0) ordering: the RCMK sweeps take minutes at this size)
    sm = gpu.Solver(Algorithm=9, Preconditioner=6, Krylov_Size=30, ILU_Ordering=1, Orthogonalization=1,
                    Precision=2, Inner_Krylov_Size=20, Nb_Iter_Max=300, Stopping_Test=1e-10)
    xm = gpu.Vector(fe.N)
    im = gpu.solve(A, bv, sm, xm)
    err_m = np.abs(xm.get()[::st] - xr).max() / scale
    print(f"C2 mixed (FP32 inner, FP64 outer): its {im['its']} rel_res {im['rel_res']:.3e}; "
          f"error against the reference's FP64 solution {err_m:.2e}", flush=True)
    assert im["rel_res"] <= 1e-5


def test_c3_batch_512(gpu, oracle):
    from getdp_amd import batch as BT
    print("case C3 batch 512", flush=True)
    probs = BT.make_instances(512)
    bt = BT.Batch(probs)
    bt.run()
    inf = bt.info()
    rel = bt.true_rel_residuals()
    print(f"C3: {int((inf['status'] == 0).sum())} of 512 converged, max true rel_res {rel.max():.3e}, "
          f"its max {int(inf['its'].max())}", flush=True)
    assert (inf["status"] == 0).all()
    # the stop test is Sparskit's Givens estimate against |r0| (F5), so the
    # true residual sits at the tolerance up to rounding
    assert (rel <= 2e-10).all(), rel.max()
    # eight instances over the size range against the oracle's solve
    sizes = np.array([p.meta["n"] for p in probs])
    cost = (sizes + 1) ** 2 * inf["its"]
    cheap = np.flatnonzero(cost <= np.percentile(cost, 60))
    order = cheap[np.argsort(sizes[cheap], kind="stable")]
    pick = order[np.linspace(0, order.size - 1, 8).astype(int)]
    assert np.unique(pick).size == 8
    for i in pick:
        ia, ja, a, b, tab = oracle.assemble(probs[i])
        xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30, maxits=20000)
        err = np.abs(bt.solution(i) - xo).max() / np.abs(xo).max()
        print(f"C3 instance {i} (n={sizes[i]}): its {int(inf['its'][i])} oracle {io['its']} err {err:.2e}",
              flush=True)
        assert err <= 1e-9
