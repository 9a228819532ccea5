"""Gmsh MSH 2.x files through the library's reader and writer
(gdp_mesh_read / gdp_mesh_write; GeoData.cpp:573-1028 semantics)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_msh(path) -> dict:
    """Nodes (xyz in ascending node number, node_num) and elements (sorted by
    number: elem_num, etype, region, elementary, conn as 1-based node
    indices, 4 per element, zero-padded)."""
    h = C.c_void_p()
    check(lib().gdp_mesh_read(str(path).encode(), C.byref(h)), "gdp_mesh_read")
    try:
        nn, ne, npt, binary = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().gdp_mesh_info(h, C.byref(nn), C.byref(ne), C.byref(npt), C.byref(binary)), "gdp_mesh_info")
        xyz = np.zeros((nn.value, 3))
        node_num = np.zeros(nn.value, np.int32)
        elem_num, etype, region, elementary = (np.zeros(ne.value, np.int32) for _ in range(4))
        conn = np.zeros((ne.value, 4), np.int32)
        check(lib().gdp_mesh_get(h, _d(xyz), _i(node_num), _i(elem_num), _i(etype), _i(region),
                                 _i(elementary), _i(conn)), "gdp_mesh_get")
    finally:
        lib().gdp_mesh_free(h)
    return dict(xyz=xyz, node_num=node_num, elem_num=elem_num, etype=etype, region=region,
                elementary=elementary, conn=conn, npoints=npt.value, binary=bool(binary.value))


def write_msh(path, xyz, etype, region, conn, binary: bool = False):
    xyz = np.ascontiguousarray(xyz, np.float64)
    etype = np.ascontiguousarray(etype, np.int32)
    region = np.ascontiguousarray(region, np.int32)
    conn = np.ascontiguousarray(conn, np.int32)
    check(lib().gdp_mesh_write(str(path).encode(), int(binary), xyz.shape[0], _d(xyz), etype.shape[0],
                               _i(etype), _i(region), _i(conn)), "gdp_mesh_write")
