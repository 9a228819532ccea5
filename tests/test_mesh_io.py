"""Gmsh MSH 2.x reader/writer (gdp_mesh_read / gdp_mesh_write) against
GeoData.cpp:573-1028's reading rules.

* tests/golden/magnet.msh is the reference's own example mesh
  (examples/magnet.msh, MSH 2.2 ASCII); an independent line-by-line Python
  parse of the same file (Python's float() is correctly rounded, as
  fscanf("%lf") is) gives the expected arrays;
* ASCII and binary round trips are exact;
* the reader's rules: elements sorted by number, region = first tag (1
  without tags in ASCII files), node numbers mapped in ascending order,
  point elements counted and dropped, unsupported versions and types
  rejected."""
from pathlib import Path

import numpy as np
import pytest

from getdp_amd import meshio
from getdp_amd._lib import GdpError

MAGNET = Path(__file__).resolve().parent / "golden" / "magnet.msh"
NNODES = {15: 1, 1: 2, 2: 3, 4: 4}


def _parse_python(path):
    lines = Path(path).read_text().splitlines()
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    nodes = {}
    for ln in lines[i + 2: i + 2 + n]:
        f = ln.split()
        nodes[int(f[0])] = [float(f[1]), float(f[2]), float(f[3])]
    i = lines.index("$Elements")
    m = int(lines[i + 1])
    els = []
    for ln in lines[i + 2: i + 2 + m]:
        f = [int(v) for v in ln.split()]
        num, typ, ntag = f[0], f[1], f[2]
        tags = f[3: 3 + ntag]
        conn = f[3 + ntag:]
        assert len(conn) == NNODES[typ]
        els.append((num, typ, tags[0] if tags else 1, conn))
    els.sort(key=lambda e: e[0])
    return nodes, els


def test_reads_the_reference_example_mesh():
    m = meshio.read_msh(MAGNET)
    nodes, els = _parse_python(MAGNET)
    ids = sorted(nodes)
    assert np.array_equal(m["node_num"], ids)
    want_xyz = np.array([nodes[k] for k in ids])
    assert np.array_equal(m["xyz"].view(np.uint64), want_xyz.view(np.uint64))
    kept = [e for e in els if e[1] != 15]
    assert m["npoints"] == len(els) - len(kept)
    assert np.array_equal(m["elem_num"], [e[0] for e in kept])
    assert np.array_equal(m["etype"], [e[1] for e in kept])
    assert np.array_equal(m["region"], [e[2] for e in kept])
    pos = {k: i + 1 for i, k in enumerate(ids)}
    conn = np.zeros((len(kept), 4), np.int32)
    for r, e in enumerate(kept):
        conn[r, : len(e[3])] = [pos[v] for v in e[3]]
    assert np.array_equal(m["conn"], conn)
    assert set(m["region"]) == {100, 101, 102, 103, 104, 105, 106}
    assert m["xyz"].shape[0] == 1691


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_is_exact(tmp_path, binary):
    m = meshio.read_msh(MAGNET)
    p = tmp_path / "m.msh"
    meshio.write_msh(p, m["xyz"], m["etype"], m["region"], m["conn"], binary=binary)
    r = meshio.read_msh(p)
    assert r["binary"] == binary
    assert np.array_equal(r["xyz"].view(np.uint64), m["xyz"].view(np.uint64))
    for k in ("etype", "region", "conn"):
        assert np.array_equal(r[k], m[k]), k
    assert np.array_equal(r["elem_num"], np.arange(1, m["etype"].shape[0] + 1))


def test_reading_rules(tmp_path):
    p = tmp_path / "t.msh"
    p.write_text("""$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
5
40 0 0 1
10 0 0 0
30 0 1 0
20 1 0 0
50 0.1 0.1 0.1
$EndNodes
$Elements
4
7 4 2 3 9 10 20 30 40
2 2 0 10 20 30
5 15 2 1 1 50
3 1 1 8 10 20
$EndElements
""")
    m = meshio.read_msh(p)
    assert np.array_equal(m["node_num"], [10, 20, 30, 40, 50])
    assert np.array_equal(m["elem_num"], [2, 3, 7])          # sorted by number, point dropped
    assert np.array_equal(m["etype"], [2, 1, 4])
    assert np.array_equal(m["region"], [1, 8, 3])            # no tags -> 1; first tag otherwise
    assert m["npoints"] == 1
    assert np.array_equal(m["conn"], [[1, 2, 3, 0], [1, 2, 0, 0], [1, 2, 3, 4]])
    assert np.array_equal(m["xyz"][3], [0, 0, 1])


@pytest.mark.parametrize("body", [
    "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n",
    "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n"
    "$Elements\n1\n1 3 2 1 1 1 1 1 1\n$EndElements\n",
    "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$EndNodes\n"
    "$Elements\n1\n1 1 2 1 1 1 2\n$EndElements\n",
])
def test_rejected_files(tmp_path, body):
    p = tmp_path / "bad.msh"
    p.write_text(body)
    with pytest.raises(GdpError):
        meshio.read_msh(p)
