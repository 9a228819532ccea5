"""Host-side pre-processing entry points of libgdp_hip (no device work)."""
from __future__ import annotations

import numpy as np

from ._lib import check, iptr, lib


def number_edges(etype: np.ndarray, conn: np.ndarray, visit: np.ndarray):
    """Signed edge numbers per element (6 per element) and edge node pairs,
    numbering edges by first encounter in ``visit`` order
    (GeoEntity.cpp:437-490)."""
    etype = np.ascontiguousarray(etype, np.int32)
    conn = np.ascontiguousarray(conn, np.int32)
    visit = np.ascontiguousarray(visit, np.int32)
    ne = etype.shape[0]
    edges = np.zeros((ne, 6), np.int32)
    max_edges = 6 * ne + 16
    en = np.zeros((max_edges, 2), np.int32)
    n = check(lib().gdp_number_edges(ne, iptr(etype), iptr(conn), visit.shape[0], iptr(visit),
                                     iptr(edges), iptr(en), max_edges), "gdp_number_edges")
    return edges, en[:n].copy()


def rcmk(ia: np.ndarray, ja: np.ndarray):
    """RCMK permutation of a CSR pattern: (perm[old]=new, rperm[new]=old, nlev)."""
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    n = ia.shape[0] - 1
    perm = np.zeros(n, np.int32)
    rperm = np.zeros(n, np.int32)
    nlev = check(lib().gdp_rcmk(n, iptr(ia), iptr(ja), iptr(perm), iptr(rperm)), "gdp_rcmk")
    return perm, rperm, nlev
