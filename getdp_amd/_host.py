"""Host-side pre-processing entry points of libgdp_hip (no device work)."""
from __future__ import annotations

import numpy as np

import ctypes as C

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


def ilut(ia: np.ndarray, ja: np.ndarray, a: np.ndarray, lfil: int = 20, droptol: float = 0.0,
         permtol: float = 0.05, pivot: bool = True, f64: bool = False):
    """ILUT / ILUTP factors as gdp_solve computes them for Preconditioner 1 / 2
    (ilut.F:52-806): (ptr, dpos, col, val, newpos), rows laid out as
    [L entries][inverted pivot][U entries]."""
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    n = ia.shape[0] - 1
    ptr = np.zeros(n + 1, np.int32)
    dpos = np.zeros(max(n, 1), np.int32)
    newpos = np.zeros(max(n, 1), np.int32)
    nz = C.c_long(0)
    dp = C.POINTER(C.c_double)
    cap = 2 * (n + 1) * (lfil + 1)  # the reference's first guess (Sparskit.cpp:314)
    for _ in range(2):
        col = np.zeros(max(cap, 1), np.int32)
        val = np.zeros(max(cap, 1), np.float64)
        r = lib().gdp_ilut(n, iptr(ia), iptr(ja), a.ctypes.data_as(dp), lfil, droptol, permtol, int(pivot),
                           int(f64), cap, iptr(ptr), iptr(dpos), iptr(col), val.ctypes.data_as(dp),
                           iptr(newpos), C.byref(nz))
        if r == 0:
            return ptr, dpos[:n], col[:nz.value], val[:nz.value], newpos[:n]
        if nz.value <= cap:
            check(r, "gdp_ilut")
        cap = nz.value
    raise RuntimeError("gdp_ilut: capacity")
