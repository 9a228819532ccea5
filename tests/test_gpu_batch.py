"""Batched assembly + ILU(0)-GMRES against the oracle, instance by instance.

* every diagonal block of the batch-assembled matrix equals the oracle's
  assembly of that instance alone (bit-exact), and so do the right-hand sides;
* every instance's solution matches the oracle's solve_matrix restatement
  (ILU(0)-GMRES(30), RCMK, FP32 factors) within 1e-9 relative, with the true
  FP64 residual <= 2e-10 relative (the stop test uses the Givens estimate, as
  the reference does), for every launch configuration the autotuner picks from;
* replays of the captured graph are bit-identical."""
import numpy as np
import pytest

from getdp_amd import batch as BT
from getdp_amd import problems as P
from getdp_amd._lib import GdpError

pytestmark = pytest.mark.gpu

SIZES = [6, 11, 8, 14, 9, 7, 12]


def small_instances():
    return [P.poisson2d(n, seed=40 + i, nblocks=4, field_kind="checker") for i, n in enumerate(SIZES)]


@pytest.fixture(scope="module")
def probs():
    return small_instances()


@pytest.fixture(scope="module")
def reference(oracle, probs):
    out = []
    for p in probs:
        ia, ja, a, b, tab = oracle.assemble(p)
        x, info, _ = oracle.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30,
                                  renumbering=1, maxits=20000)
        out.append((ia, ja, a, b, x, info))
    return out


def test_batch_assembly_blocks_equal_single_assembly(gpu, probs, reference):
    bt = BT.Batch(probs, config=[BT.DEFAULT_CFG] * len(probs))
    bt.run()
    ia, ja, a = bt.A.get_csr()
    b = bt.b.get()
    for i, (ria, rja, ra, rb, _, _) in enumerate(reference):
        r0, r1 = bt.off[i], bt.off[i + 1]
        assert r1 - r0 == ria.shape[0] - 1
        sia = ia[r0:r1 + 1] - ia[r0]
        sja = ja[ia[r0]:ia[r1]] - r0
        sa = a[ia[r0]:ia[r1]]
        assert np.array_equal(sia, ria) and np.array_equal(sja, rja)
        assert np.array_equal(sa, ra)  # bit-exact
        assert np.array_equal(b[r0:r1], rb)


@pytest.mark.parametrize("cfg", BT.CANDIDATES)
def test_batch_solutions_match_oracle(gpu, probs, reference, cfg):
    print(f"case batch cfg {cfg}", flush=True)
    bt = BT.Batch(probs, config=[cfg] * len(probs))
    bt.run()
    inf = bt.info()
    assert (inf["status"] == 0).all(), inf["status"]
    rel = bt.true_rel_residuals()
    assert (rel <= 2e-10).all(), rel
    for i, (_, _, _, _, xo, io) in enumerate(reference):
        xs = bt.solution(i)
        err = np.abs(xs - xo).max() / np.abs(xo).max()
        assert err <= 1e-9, (i, err)
        if cfg[2] == 0:
            # same ordering and factors as the reference: comparable counts
            assert abs(int(inf["its"][i]) - io["its"]) <= max(3, io["its"] // 10), (inf["its"][i], io["its"])


def test_mixed_configurations_and_replay(gpu, probs, reference):
    cfgs = [BT.CANDIDATES[i % len(BT.CANDIDATES)] for i in range(len(probs))]
    bt = BT.Batch(probs, config=cfgs)
    bt.run()
    x1 = bt.x.get()
    bt.run()
    x2 = bt.x.get()
    assert np.array_equal(x1, x2)
    inf = bt.info()
    assert (inf["status"] == 0).all()
    assert (inf["t_end_us"] >= inf["t_begin_us"]).all()
    for i, (_, _, _, _, xo, _) in enumerate(reference):
        assert np.abs(bt.solution(i) - xo).max() <= 1e-9 * np.abs(xo).max()


def test_iteration_limit_reports_status(gpu, probs):
    bt = BT.Batch(probs, config=[BT.DEFAULT_CFG] * len(probs), maxits=3)
    bt.run()
    inf = bt.info()
    assert (inf["status"] == -1).all()
    assert (inf["its"] == 3).all()


def test_batch_rejects_unsupported_parameters(gpu, probs):
    bt = BT.Batch(probs[:2])
    bad = gpu.Solver(Algorithm=1, Preconditioner=6)
    import ctypes as C
    from getdp_amd._lib import iptr, lib
    h = C.c_void_p()
    rc = lib().gdp_batch_create(bad.h, 2, iptr(bt.off), None, C.byref(h))
    assert rc == -4  # GDP_E_UNSUPPORTED


@pytest.mark.parametrize("walk", ["1", "0"])
@pytest.mark.parametrize("threads", [256, 512, 1024])
def test_large_instances_lds_ring(gpu, oracle, threads, walk, monkeypatch):
    """Instances larger than the LDS sweep vector (4096, 8192, 14336 rows at
    256, 512, 1024 threads) sweep through the LDS ring of recent results: the
    solutions match the oracle's, with the reference's iteration counts (same
    ordering and factors), next to small instances on the LDS path; sweeps
    walked by one wave (the default) or level by level (GDP_BATCH_WALK=0)."""
    monkeypatch.setenv("GDP_BATCH_WALK", walk)
    sizes = [70, 150, 12, 96, 135]  # 71^2 = 5041 ... 151^2 = 22801 rows
    probs = [P.poisson2d(n, seed=70 + i, nblocks=8, field_kind="checker") for i, n in enumerate(sizes)]
    print(f"case batch ring threads {threads} walk {walk}", flush=True)
    bt = BT.Batch(probs, config=[(threads, 0, 0)] * len(probs))
    bt.run()
    inf = bt.info()
    assert (inf["status"] == 0).all(), inf["status"]
    assert (bt.true_rel_residuals() <= 2e-10).all()
    for i, p in enumerate(probs):
        ia, ja, a, b, tab = oracle.assemble(p)
        xo, io, _ = oracle.solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30, renumbering=1,
                                 maxits=20000)
        err = np.abs(bt.solution(i) - xo).max() / np.abs(xo).max()
        assert err <= 1e-9, (i, err)
        assert abs(int(inf["its"][i]) - io["its"]) <= max(3, io["its"] // 10), (i, inf["its"][i], io["its"])


def test_walked_and_level_sweeps_identical(gpu, monkeypatch):
    """The one-wave walk and the level-by-level sweeps apply the factors with
    the same operations per row: the batch gives the same iteration counts and
    the same solution bits either way, for instances on the LDS path and on
    the LDS-ring path."""
    sizes = [40, 140, 20, 100]
    probs = [P.poisson2d(n, seed=90 + i, nblocks=8, field_kind="checker") for i, n in enumerate(sizes)]
    out = {}
    for walk in ("1", "0"):
        monkeypatch.setenv("GDP_BATCH_WALK", walk)
        bt = BT.Batch(probs, config=[(1024, 0, 0)] * len(probs))
        bt.run()
        out[walk] = (bt.info()["its"].copy(), bt.x.get().copy())
        bt.close()
    assert np.array_equal(out["1"][0], out["0"][0])
    assert np.array_equal(out["1"][1].view(np.uint64), out["0"][1].view(np.uint64))
