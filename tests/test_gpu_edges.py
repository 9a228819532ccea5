"""Device edge numbering (sort-based first-encounter numbering) against the
oracle's sequential walk (orc_number_edges, GeoEntity.cpp:437-490), which the
CPU tests pin on hand-worked cases.  Bar: identical signed numbers and node
pairs."""
import numpy as np
import pytest

from getdp_amd import _host, problems as P

pytestmark = pytest.mark.gpu


def _check(oracle, etype, conn, visit):
    e1, n1 = _host.number_edges(etype, conn, visit)      # device (library initialised)
    e2, n2 = oracle.number_edges(etype, conn, visit)
    assert np.array_equal(e1, e2)
    assert np.array_equal(n1, n2)
    return n1.shape[0]


@pytest.mark.parametrize("n", [2, 5, 24])
def test_curlcurl_mesh_edges(gpu, oracle, n):
    pr = P.curlcurl3d(n)
    ne = _check(oracle, pr.etype, pr.conn, pr.meta["visit"])
    assert ne == pr.edge_nodes.shape[0]


def test_mixed_types_repeats_and_partial_visits(gpu, oracle):
    """Lines, triangles and tetrahedra; elements visited twice (the repeat is
    skipped) and elements never visited (zero rows)."""
    pr = P.curlcurl3d(3)
    rng = np.random.default_rng(9)
    etype = pr.etype.copy()
    conn = pr.conn.copy()
    lines = rng.choice(np.where(etype == P.TRI)[0], 20, replace=False)
    etype[lines] = P.LINE
    ne = etype.shape[0]
    visit = rng.permutation(ne)[: ne - 50]
    visit = np.concatenate([visit, visit[:100]])  # repeats
    visit = rng.permutation(visit).astype(np.int32)
    _check(oracle, etype, conn, visit)


def test_device_numbering_errors(gpu):
    pr = P.curlcurl3d(2)
    from getdp_amd._lib import GdpError
    bad = pr.meta["visit"].copy()
    bad[3] = pr.etype.shape[0] + 5
    with pytest.raises(GdpError):
        _host.number_edges(pr.etype, pr.conn, bad)
    et = pr.etype.copy()
    et[0] = 7  # unsupported element type
    with pytest.raises(GdpError):
        _host.number_edges(et, pr.conn, pr.meta["visit"])
