"""Synthetic finite-element problems: structured meshes, function spaces and
Galerkin terms, described the way GetDP's .pro front-end would hand them to
the assembly (elements sorted by number, per-region coefficients, constrained
entities).

The mesh numbering follows Gmsh/GetDP conventions: node and element numbers
are 1-based, elements are processed in ascending number
(src/kernel/GeoData.cpp:1018), element types use Gmsh codes (2 triangle,
4 tetrahedron, 1 line).

Problems built here:
  * poisson2d(n)      2D scalar magnetostatics/Poisson, P1 triangles
                      (``[ nu[] * Dof{d u}, {d u} ] + [ -js[], {u} ]``)
  * elasticity3d(n)   3D linear elasticity, 6-tet Kuhn cubes, the tensor form
                      of templates/Lib_Elasticity_u.pro:151-219 (block DOFs)
  * poisson3d(n)      3D scalar Poisson, P1 on 6-tet Kuhn cells, checkerboard
                      or lognormal nu (the learned preconditioner's 3D problem)
  * curlcurl3d(n)     3D magnetostatics with Whitney edge elements
                      (``[ nu[] * Dof{d a}, {d a} ] + [ 1e-6 * Dof{a}, {a} ]
                      + [ -js[], {a} ]``), PEC boundary
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# Element types (Gmsh codes)
LINE, TRI, TET = 1, 2, 4
# Basis function kinds
BF_NODE, BF_EDGE, BF_PERP_EDGE = 0, 1, 2
# Term kinds (shared with include/gdp_hip.h and the oracle)
T_GRAD_GRAD, T_CURL_CURL, T_EDGE_MASS, T_NODE_MASS = 1, 2, 3, 4
T_SRC_NODE, T_SRC_EDGE = 10, 11


@dataclass
class Space:
    """A FunctionSpace with one basis function (BF_Node, BF_Edge or BF_PerpendicularEdge)."""
    bf: int
    numtype: int                       # BasisFunction.Num: the DOF key's first field
    support: List[int]
    fixed_entity: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    fixed_value: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))


@dataclass
class Term:
    """One Galerkin term ``Integral { [ coef * Dof{op v}, {op u} ]; In regions; }``."""
    kind: int
    equ: int                 # test-function space index
    dof: int                 # unknown space index, -1 for source terms
    regions: List[int]
    npoints: int             # Gauss points
    coef: np.ndarray         # (nregion_ids, ncoef) whole coefficient, sign included
    jac_sur: int = 0
    dim: int = 3

    @property
    def ncoef(self) -> int:
        return int(self.coef.shape[1])


@dataclass
class Problem:
    name: str
    xyz: np.ndarray          # (Nn, 3) float64
    etype: np.ndarray        # (E,) int32
    region: np.ndarray       # (E,) int32
    conn: np.ndarray         # (E, 4) int32, 1-based node numbers
    spaces: List[Space]
    terms: List[Term]
    edges: Optional[np.ndarray] = None       # (E, 6) int32 signed edge numbers
    edge_nodes: Optional[np.ndarray] = None  # (nedges, 2)
    grid: Optional[tuple] = None             # structured-grid shape (for the U-Net maps)
    meta: dict = field(default_factory=dict)

    @property
    def nnodes(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.etype.shape[0])

    def nmax_entity(self) -> int:
        ne = 0 if self.edge_nodes is None else int(self.edge_nodes.shape[0])
        return max(self.nnodes, ne)


# ---------------------------------------------------------------------------
# 2D unit square, two P1 triangles per cell


def _square_mesh(n: int):
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    xyz = np.zeros(((n + 1) ** 2, 3))
    xyz[:, 0] = (i.ravel() / n)
    xyz[:, 1] = (j.ravel() / n)

    def nid(ii, jj):
        return jj * (n + 1) + ii + 1

    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    n00, n10, n11, n01 = nid(ci, cj), nid(ci + 1, cj), nid(ci + 1, cj + 1), nid(ci, cj + 1)
    tri = np.zeros((2 * n * n, 4), np.int32)
    tri[0::2, :3] = np.stack([n00, n10, n11], 1)
    tri[1::2, :3] = np.stack([n00, n11, n01], 1)
    cell_i = np.repeat(ci, 2)
    cell_j = np.repeat(cj, 2)
    k = np.arange(n)
    lines = np.concatenate([
        np.stack([nid(k, 0), nid(k + 1, 0)], 1),
        np.stack([nid(n, k), nid(n, k + 1)], 1),
        np.stack([nid(n - k, n), nid(n - k - 1, n)], 1),
        np.stack([nid(0, n - k), nid(0, n - k - 1)], 1),
    ]).astype(np.int32)
    return xyz, tri, cell_i, cell_j, lines


def poisson2d(n: int = 256, seed: int = 0, nblocks: int = 16, nu_vals=(1.0, 1e-3),
              js: float = 1.0, radius: float = 0.2, field_kind: str = "checker",
              sigma: float = 1.0, nbins: int = 64, npoints: int = 3,
              basis: str = "node") -> Problem:
    """2D scalar Poisson/magnetostatics on the unit square.

    Regions: triangles get ``1 + cls + K*in_src`` where cls is the coefficient
    class (checkerboard: 2 classes; lognormal: ``nbins`` quantised classes),
    boundary lines region 0.

    ``basis="node"``: u in BF_Node, ``[ nu * Dof{d u}, {d u} ] + [ -js, {u} ]``.
    ``basis="perp"``: the magnetostatic a_z form of the same problem, a in
    BF_PerpendicularEdge, ``[ nu * Dof{d a}, {d a} ] + [ -Vector[0,0,js], {a} ]``
    (curl (phi e_z) . curl (psi e_z) = grad phi . grad psi: the same system up
    to rounding).
    """
    xyz, tri, ci, cj, lines = _square_mesh(n)
    ntri = tri.shape[0]
    rng = np.random.default_rng(seed)
    if field_kind == "checker":
        blocks = rng.integers(0, 2, size=(nblocks, nblocks))
        bx = np.minimum(ci * nblocks // n, nblocks - 1)
        by = np.minimum(cj * nblocks // n, nblocks - 1)
        cls = blocks[by, bx].astype(np.int32)
        K = 2
        cls_val = np.asarray(nu_vals, np.float64)
    else:  # per-cell lognormal, quantised to nbins log-uniform classes
        z = rng.normal(0.0, sigma, size=n * n)
        edges_ = np.linspace(-4 * sigma, 4 * sigma, nbins + 1)
        b = np.clip(np.searchsorted(edges_, z) - 1, 0, nbins - 1)
        cls = np.repeat(b, 2).astype(np.int32)
        K = nbins
        cls_val = np.exp(0.5 * (edges_[:-1] + edges_[1:]))
    cen = xyz[tri[:, :3] - 1].mean(axis=1)
    src = (((cen[:, 0] - 0.5) ** 2 + (cen[:, 1] - 0.5) ** 2) < radius ** 2).astype(np.int32)
    reg_tri = 1 + cls + K * src
    nreg = 1 + 2 * K
    etype = np.concatenate([np.full(ntri, TRI), np.full(lines.shape[0], LINE)]).astype(np.int32)
    region = np.concatenate([reg_tri, np.zeros(lines.shape[0], np.int32)]).astype(np.int32)
    conn = np.zeros((etype.shape[0], 4), np.int32)
    conn[:ntri] = tri
    conn[ntri:, :2] = lines
    dom = list(range(1, nreg))
    srcreg = list(range(1 + K, nreg))
    nu = np.zeros((nreg, 1))
    for r in range(1, nreg):
        nu[r, 0] = cls_val[(r - 1) % K]
    f = np.zeros((nreg, 1))
    f[srcreg, 0] = -js
    bnodes = np.unique(lines.ravel()).astype(np.int32)
    if basis == "perp":
        space = Space(BF_PERP_EDGE, 1, dom, bnodes, np.zeros(bnodes.shape[0]))
        f3 = np.zeros((nreg, 3))
        f3[:, 2] = f[:, 0]
        terms = [Term(T_CURL_CURL, 0, 0, dom, npoints, nu, dim=2),
                 Term(T_SRC_EDGE, 0, -1, srcreg, npoints, f3, dim=2)]
    elif basis == "node":
        space = Space(BF_NODE, 1, dom, bnodes, np.zeros(bnodes.shape[0]))
        terms = [Term(T_GRAD_GRAD, 0, 0, dom, npoints, nu, dim=2),
                 Term(T_SRC_NODE, 0, -1, srcreg, npoints, f, dim=2)]
    else:
        raise ValueError(f"basis {basis!r}")
    return Problem(f"poisson2d_{n}" + ("_az" if basis == "perp" else ""), xyz, etype, region, conn, [space], terms,
                   grid=(n + 1, n + 1),
                   meta=dict(n=n, seed=seed, ncls=K, cls_val=cls_val, ntri=ntri))


# ---------------------------------------------------------------------------
# 3D unit cube, six-tetrahedron Kuhn split of every hexahedral cell

_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _cube_mesh(n: int):
    r = np.arange(n + 1)
    kk, jj, ii = np.meshgrid(r, r, r, indexing="ij")
    xyz = np.stack([ii.ravel() / n, jj.ravel() / n, kk.ravel() / n], 1).astype(np.float64)

    def nid(i, j, k):
        return (k * (n + 1) + j) * (n + 1) + i + 1

    c = np.arange(n)
    ck, cj, ci = np.meshgrid(c, c, c, indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    base = np.stack([ci, cj, ck], 1)
    tets = np.zeros((6 * n ** 3, 4), np.int32)
    for t, perm in enumerate(_KUHN):
        p = base.copy()
        v = [nid(p[:, 0], p[:, 1], p[:, 2])]
        for ax in perm[:2]:
            p[:, ax] += 1
            v.append(nid(p[:, 0], p[:, 1], p[:, 2]))
        v.append(nid(ci + 1, cj + 1, ck + 1))
        tets[t::6] = np.stack(v, 1)
    cell = np.repeat(np.arange(n ** 3), 6)
    return xyz, tets, ci, cj, ck, cell, nid


def _face_tris(n: int, nid, axis: int, side: int):
    """Boundary triangles of the face x_axis = side, split along the diagonal
    that matches the Kuhn tetrahedra (corner with the lowest in-face
    coordinates to the corner with the highest)."""
    a, b = [d for d in range(3) if d != axis]
    c = np.arange(n)
    cb, ca = np.meshgrid(c, c, indexing="ij")
    ca, cb = ca.ravel(), cb.ravel()

    def node(pa, pb):
        p = [None, None, None]
        p[axis] = np.full_like(pa, side * n)
        p[a] = pa
        p[b] = pb
        return nid(p[0], p[1], p[2])

    v00, v10, v11, v01 = node(ca, cb), node(ca + 1, cb), node(ca + 1, cb + 1), node(ca, cb + 1)
    t = np.zeros((2 * n * n, 4), np.int32)
    t[0::2, :3] = np.stack([v00, v10, v11], 1)
    t[1::2, :3] = np.stack([v00, v11, v01], 1)
    return t


def lame_tensors(E: np.ndarray, nu: float):
    """-C_ij tensors of templates/Lib_Elasticity_u.pro:98-118 (3D), evaluated in
    the template's order of operations.  Returns (len(E), 9 terms, 9)."""
    E = np.asarray(E, np.float64)
    a = E * (1. - nu) / (1. + nu) / (1. - 2. * nu)
    c = E * nu / (1. + nu) / (1. - 2. * nu)
    b = E / 2. / (1. + nu)
    z = np.zeros_like(E)
    C = {
        "xx": [a, z, z, z, b, z, z, z, b], "xy": [z, c, z, b, z, z, z, z, z],
        "xz": [z, z, c, z, z, z, b, z, z], "yx": [z, b, z, c, z, z, z, z, z],
        "yy": [b, z, z, z, a, z, z, z, b], "yz": [z, z, z, z, z, c, z, b, z],
        "zx": [z, z, b, z, z, z, c, z, z], "zy": [z, z, z, z, z, b, z, c, z],
        "zz": [b, z, z, z, b, z, z, z, a],
    }
    order = ["xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz"]
    return np.stack([-np.stack(C[k], 1) for k in order], 1)


def elasticity3d(n: int = 64, seed: int = 0, sigma: float = 1.0, poisson: float = 0.3,
                 nbins: int = 64, traction: float = 1.0, npoints: int = 4) -> Problem:
    """3D linear elasticity (block DOF ordering ux, uy, uz; F7 contract)."""
    xyz, tets, ci, cj, ck, cell, nid = _cube_mesh(n)
    nt = tets.shape[0]
    rng = np.random.default_rng(seed)
    z = rng.normal(0.0, sigma, size=nt)
    edges_ = np.linspace(-4 * sigma, 4 * sigma, nbins + 1)
    b = np.clip(np.searchsorted(edges_, z) - 1, 0, nbins - 1).astype(np.int32)
    Evals = np.exp(0.5 * (edges_[:-1] + edges_[1:]))
    trac = _face_tris(n, nid, 0, 1)
    RTRAC = nbins
    etype = np.concatenate([np.full(nt, TET), np.full(trac.shape[0], TRI)]).astype(np.int32)
    region = np.concatenate([b, np.full(trac.shape[0], RTRAC)]).astype(np.int32)
    conn = np.concatenate([tets, trac]).astype(np.int32)
    nreg = nbins + 1
    Ct = lame_tensors(Evals, poisson)            # (nbins, 9, 9)
    dom = list(range(nbins))
    clamped = (np.flatnonzero(xyz[:, 0] == 0.0) + 1).astype(np.int32)
    # the displacement spaces' support covers the loaded boundary as well
    # (Dom_Mec = Vol_Mec + Sur_Neu_Mec in the elasticity template)
    support = dom + [RTRAC]
    spaces = [Space(BF_NODE, s + 1, support, clamped, np.zeros(clamped.shape[0]))
              for s in range(3)]
    terms = []
    for t9 in range(9):
        coef = np.zeros((nreg, 9))
        coef[:nbins] = Ct[:, t9, :]
        terms.append(Term(T_GRAD_GRAD, t9 // 3, t9 % 3, dom, npoints, coef))
    f = np.zeros((nreg, 1))
    f[RTRAC, 0] = traction
    terms.append(Term(T_SRC_NODE, 0, -1, [RTRAC], 3, f, jac_sur=1))
    return Problem(f"elasticity3d_{n}", xyz, etype, region, conn, spaces, terms,
                   meta=dict(n=n, seed=seed, E=Evals, ntet=nt))


def poisson3d(n: int = 64, seed: int = 0, nsub: int = 8, nu_vals=(1.0, 1e-3), frac: float = 0.5,
              field_kind: str = "checker", sigma: float = 1.0, nbins: int = 64, js: float = 1.0,
              radius: float = 0.25, npoints: int = 4) -> Problem:
    """3D scalar magnetostatics/Poisson on the unit cube, 6-tet Kuhn cells:
    ``[ nu[] * Dof{d u}, {d u} ] + [ -js[], {u} ]`` with P1 BF_Node, u = 0 on
    the whole boundary.  nu per element: ``checker`` -- a seeded random
    ``frac`` of the nsub^3 subcubes take nu_vals[1], the rest nu_vals[0];
    ``lognormal`` -- per-cell lognormal (sigma), quantised to nbins classes.
    Source js inside a centred ball of the given radius.

    Regions: tetrahedra ``1 + cls + K*in_src`` (K classes), boundary triangles 0."""
    xyz, tets, ci, cj, ck, cell, nid = _cube_mesh(n)
    nt = tets.shape[0]
    rng = np.random.default_rng(seed)
    if field_kind == "checker":
        nsc = nsub ** 3
        chosen = np.zeros(nsc, np.int32)
        chosen[rng.choice(nsc, size=int(round(frac * nsc)), replace=False)] = 1
        sc = ((ck * nsub // n) * nsub + (cj * nsub // n)) * nsub + (ci * nsub // n)
        cls = np.repeat(chosen[sc], 6).astype(np.int32)
        K = 2
        cls_val = np.asarray(nu_vals, np.float64)
    else:
        z = rng.normal(0.0, sigma, size=n ** 3)
        edges_ = np.linspace(-4 * sigma, 4 * sigma, nbins + 1)
        cls = np.repeat(np.clip(np.searchsorted(edges_, z) - 1, 0, nbins - 1), 6).astype(np.int32)
        K = nbins
        cls_val = np.exp(0.5 * (edges_[:-1] + edges_[1:]))
    cen = xyz[tets - 1].mean(axis=1)
    src = (((cen - 0.5) ** 2).sum(axis=1) < radius ** 2).astype(np.int32)
    reg_tet = 1 + cls + K * src
    nreg = 1 + 2 * K
    bnd = np.concatenate([_face_tris(n, nid, ax, sd) for ax in range(3) for sd in (0, 1)])
    etype = np.concatenate([np.full(nt, TET), np.full(bnd.shape[0], TRI)]).astype(np.int32)
    region = np.concatenate([reg_tet, np.zeros(bnd.shape[0], np.int32)]).astype(np.int32)
    conn = np.concatenate([tets, bnd]).astype(np.int32)
    dom = list(range(1, nreg))
    srcreg = list(range(1 + K, nreg))
    nu = np.zeros((nreg, 1))
    for r in range(1, nreg):
        nu[r, 0] = cls_val[(r - 1) % K]
    f = np.zeros((nreg, 1))
    f[srcreg, 0] = -js
    bnodes = np.unique(bnd[:, :3].ravel()).astype(np.int32)
    space = Space(BF_NODE, 1, dom, bnodes, np.zeros(bnodes.shape[0]))
    terms = [Term(T_GRAD_GRAD, 0, 0, dom, npoints, nu),
             Term(T_SRC_NODE, 0, -1, srcreg, npoints, f)]
    return Problem(f"poisson3d_{n}", xyz, etype, region, conn, [space], terms,
                   grid=(n + 1, n + 1, n + 1),
                   meta=dict(n=n, seed=seed, ncls=K, cls_val=cls_val, ntet=nt))


def curlcurl3d(n: int = 48, seed: int = 0, nsub: int = 8, frac: float = 0.2,
               nu_vals=(1.0, 1e-3), radius: float = 0.25, mass: float = 1e-6,
               npoints: int = 4) -> Problem:
    """3D curl-curl with Whitney edges, PEC on the whole boundary."""
    from . import _host  # edge numbering lives in the native host library

    xyz, tets, ci, cj, ck, cell, nid = _cube_mesh(n)
    nt = tets.shape[0]
    rng = np.random.default_rng(seed)
    nsc = nsub ** 3
    chosen = np.zeros(nsc, np.int32)
    chosen[rng.choice(nsc, size=int(round(frac * nsc)), replace=False)] = 1
    sc = ((ck * nsub // n) * nsub + (cj * nsub // n)) * nsub + (ci * nsub // n)
    cls = np.repeat(chosen[sc], 6)
    cen = xyz[tets[:, :4] - 1].mean(axis=1)
    src = (((cen[:, 0] - 0.5) ** 2 + (cen[:, 1] - 0.5) ** 2) < radius ** 2).astype(np.int32)
    reg_tet = 1 + cls + 2 * src
    bnd = np.concatenate([_face_tris(n, nid, ax, s) for ax in range(3) for s in (0, 1)])
    etype = np.concatenate([np.full(nt, TET), np.full(bnd.shape[0], TRI)]).astype(np.int32)
    region = np.concatenate([reg_tet, np.zeros(bnd.shape[0], np.int32)]).astype(np.int32)
    conn = np.concatenate([tets, bnd]).astype(np.int32)
    # Edge creation order: the first volume element's edges are created when
    # its DOFs are gathered (Get_DofOfElement.cpp:120-124), whose constraint
    # check then builds the EdgesOf[boundary] group, numbering the boundary
    # triangles' edges in element order (ExtendedGroup.cpp:210-233), before
    # the remaining volume elements.
    visit = np.concatenate([[0], np.arange(nt, nt + bnd.shape[0]), np.arange(1, nt)]).astype(np.int32)
    edges, edge_nodes = _host.number_edges(etype, conn, visit)
    bedges = np.unique(np.abs(edges[nt:, :3]).ravel()).astype(np.int32)
    dom = [1, 2, 3, 4]
    nu = np.zeros((5, 1))
    nu[[1, 3], 0] = nu_vals[0]
    nu[[2, 4], 0] = nu_vals[1]
    ms = np.zeros((5, 1))
    ms[1:, 0] = mass
    f = np.zeros((5, 3))
    f[[3, 4], 2] = -1.0
    space = Space(BF_EDGE, 1, dom, bedges, np.zeros(bedges.shape[0]))
    terms = [Term(T_CURL_CURL, 0, 0, dom, npoints, nu),
             Term(T_EDGE_MASS, 0, 0, dom, npoints, ms),
             Term(T_SRC_EDGE, 0, -1, [3, 4], npoints, f)]
    return Problem(f"curlcurl3d_{n}", xyz, etype, region, conn, [space], terms,
                   edges=edges, edge_nodes=edge_nodes,
                   meta=dict(n=n, seed=seed, ntet=nt, visit=visit))
