"""The library's host ILUT / ILUTP factorisation (gdp_ilut, the factors
gdp_solve applies for Preconditioner 1 / 2) against the oracle's restatement
of ilut.F, which the reference's Fortran pins bit for bit
(test_oracle_sparskit.py): the same entries in the same accumulation order,
the same values bit for bit, the same column permutation."""
import numpy as np
import pytest

from getdp_amd import _host

CASES = ["poisson2d_8", "elasticity3d_2", "curlcurl3d_2"]


def _random_csr(rng, n, extra, diag_boost):
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


def _oracle_rows(o, n):
    """Oracle MSR factors in the library's row layout [L][pivot][U]."""
    alu, jlu, ju, iperm = o[0], o[1], o[2], o[3]
    ptr, col, val, dpos = [0], [], [], []
    for i in range(n):
        lo, mid, hi = jlu[i], ju[i], jlu[i + 1]
        col += list(jlu[lo:mid])
        val += list(alu[lo:mid])
        dpos.append(len(col))
        col.append(i)
        val.append(alu[i])
        col += list(jlu[mid:hi])
        val += list(alu[mid:hi])
        ptr.append(len(col))
    return (np.array(ptr), np.array(dpos), np.array(col), np.array(val, alu.dtype), iperm[n:2 * n])


def _cases(oracle, golden):
    rng = np.random.default_rng(21)
    out = [_random_csr(rng, 180, 600, boost) for boost in (1.0, 0.05)]
    for name in CASES:
        ia, ja, a = golden[f"{name}/ia"], golden[f"{name}/ja"], golden[f"{name}/a"]
        perm, rperm, ib, jb, bb, nlev = oracle.rcmk(ia, ja, a)
        out.append((ib, jb, bb))
    return out


@pytest.mark.parametrize("pivot", [False, True])
@pytest.mark.parametrize("f64", [False, True])
def test_host_ilut_matches_oracle(oracle, golden, pivot, f64):
    permuted = 0
    for ia, ja, a in _cases(oracle, golden):
        n = ia.shape[0] - 1
        for lfil, droptol, permtol in ((20, 0.0, 0.05), (5, 1e-3, 0.05), (3, 0.0, 0.5), (0, 0.0, 0.05)):
            o = oracle.ilut(ia, ja, a, lfil, droptol, permtol, pivot, f64)
            assert o[5] == 0
            want = _oracle_rows(o, n)
            got = _host.ilut(ia, ja, a, lfil, droptol, permtol, pivot, f64)
            for g, w in zip(got[:3], want[:3]):
                assert np.array_equal(g, w)
            wv = want[3].astype(np.float64)  # float factors are returned exactly in double
            assert np.array_equal(got[3].view(np.uint64), wv.view(np.uint64))
            if pivot:
                assert np.array_equal(got[4], want[4])
                permuted += int(np.any(got[4] != np.arange(n)))
            else:
                assert np.array_equal(got[4], np.arange(n))
    if pivot:
        assert permuted > 0


def test_host_ilut_sweeps_reproduce_oracle_lusol(oracle, golden):
    """(LU)^-1 y from the row layout, accumulated in storage order, equals the
    oracle's lusol on its MSR factors bit for bit (the device sweeps follow the
    same order)."""
    ia, ja, a = _cases(oracle, golden)[1]
    n = ia.shape[0] - 1
    y = np.random.default_rng(3).standard_normal(n)
    o = oracle.ilut(ia, ja, a, 20, 0.0, 0.5, True, False)
    want = oracle.lusol(o[0], o[1], o[2], y)
    ptr, dpos, col, val, newpos = _host.ilut(ia, ja, a, 20, 0.0, 0.5, True, False)
    x = np.zeros(n)
    for i in range(n):
        s = y[i]
        for k in range(ptr[i], dpos[i]):
            s = s - val[k] * x[col[k]]
        x[i] = s
    for i in range(n - 1, -1, -1):
        s = x[i]
        for k in range(dpos[i] + 1, ptr[i + 1]):
            s = s - val[k] * x[col[k]]
        x[i] = val[dpos[i]] * s
    assert np.array_equal(x, want)


def test_host_ilut_errors():
    from getdp_amd._lib import GdpError
    ia = np.array([0, 1, 2], np.int32)
    ja = np.array([0, 1], np.int32)
    with pytest.raises(GdpError, match="Zero row"):
        _host.ilut(ia, ja, np.array([1.0, 0.0]))
    with pytest.raises(GdpError, match="nb_fill"):
        _host.ilut(ia, ja, np.array([1.0, 2.0]), lfil=-1)
