"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (liboracle.so)
and to the reference's own Sparskit kernels compiled in oracle/_ref.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-C", str(HERE), "-s", "all"], check=True)


class Params(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("preconditioner", C.c_int),
                ("precond_position", C.c_int), ("krylov", C.c_int), ("maxits", C.c_int),
                ("renumbering", C.c_int), ("scaling", C.c_int), ("ilu_double", C.c_int),
                ("rtol", C.c_double), ("nb_fill", C.c_int), ("droptol", C.c_double),
                ("permtol", C.c_double)]


class Info(C.Structure):
    _fields_ = [("its", C.c_int), ("status", C.c_int), ("ilu_ierr", C.c_int),
                ("res_est", C.c_double), ("res0", C.c_double), ("true_res", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefParams(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("preconditioner", C.c_int),
                ("precond_position", C.c_int), ("krylov", C.c_int), ("maxits", C.c_int),
                ("renumbering", C.c_int), ("scaling", C.c_int), ("rtol", C.c_double),
                ("nb_fill", C.c_int), ("droptol", C.c_double), ("permtol", C.c_double)]


class RefInfo(C.Structure):
    _fields_ = [("its", C.c_int), ("status", C.c_int), ("ilu_ierr", C.c_int),
                ("res_est", C.c_double), ("res1", C.c_double), ("true_res", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OSpace(C.Structure):
    _fields_ = [("bf", C.c_int), ("numtype", C.c_int), ("nsupport", C.c_int),
                ("support", _ip), ("nfixed", C.c_int), ("fixed_entity", _ip),
                ("fixed_value", _dp)]


class OTerm(C.Structure):
    _fields_ = [("kind", C.c_int), ("equ", C.c_int), ("dof", C.c_int), ("nregions", C.c_int),
                ("regions", _ip), ("npoints", C.c_int), ("jac_sur", C.c_int), ("dim", C.c_int),
                ("ncoef", C.c_int), ("coef_nreg", C.c_int), ("coef", _dp)]


class OMesh(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("xyz", _dp), ("nelem", C.c_int), ("etype", _ip),
                ("region", _ip), ("conn", _ip), ("edges", _ip)]


def _i(a):
    return a.ctypes.data_as(_ip)


def _d(a):
    return a.ctypes.data_as(_dp)


_O = None
_R = {}


def lib():
    global _O
    if _O is None:
        p = HERE / "liboracle.so"
        if not p.exists():
            build()
        _O = C.CDLL(str(p))
        _O.orc_ddot.restype = C.c_double
        _O.orc_assemble.restype = C.c_long
    return _O


def set_dot_blocks(nblocks: int = 0):
    """Tests only: sum every dot product in ``nblocks`` contiguous blocks
    (0: the reference's order), to measure what rounding alone changes."""
    lib().orc_set_dot_blocks(int(nblocks))


def set_threads(nthreads: int = 1):
    """CPU baselines only: run the element loop, amux and the CG/GMRES vector
    loops on ``nthreads`` OpenMP threads (dot products then summed in 64
    ordered blocks); 1 restores the reference's sequential order."""
    lib().orc_set_threads(int(nthreads))


def ref(f64_ilu: bool = False):
    """The reference's Fortran Sparskit kernels (None if not built here)."""
    key = "f64" if f64_ilu else "f32"
    if key not in _R:
        p = HERE / "_ref" / ("libsparskit_ref_f64.so" if f64_ilu else "libsparskit_ref.so")
        _R[key] = C.CDLL(str(p)) if p.exists() else None
        if _R[key] is not None:
            _R[key].dnrm2_.restype = C.c_double
    return _R[key]


# ---- Sparskit restatement --------------------------------------------------

def amux(ia, ja, a, x):
    n = ia.shape[0] - 1
    y = np.zeros(n)
    lib().orc_amux(n, _d(np.ascontiguousarray(x, np.float64)), _d(y), _i(ia), _i(ja), _d(a))
    return y


def ilu0(ia, ja, a, double=False):
    n = ia.shape[0] - 1
    nz = int(ia[n]) + n + 1
    alu = np.zeros(nz, np.float64 if double else np.float32)
    jlu = np.zeros(nz, np.int32)
    ju = np.zeros(n + 1, np.int32)
    f = lib().orc_ilu0_f64 if double else lib().orc_ilu0_f32
    ptr = alu.ctypes.data_as(_dp if double else _fp)
    ierr = f(n, _i(ia), _i(ja), _d(a), ptr, _i(jlu), _i(ju))
    return alu, jlu, ju, ierr


def ilut(ia, ja, a, lfil=20, droptol=0.0, permtol=0.05, pivot=False, double=False, iwk=None):
    """Oracle ilut / ilutp: (alu, jlu, ju, iperm, ja_after, ierr), 0-based MSR."""
    n = ia.shape[0] - 1
    iwk = iwk or 2 * (n + 1) * (lfil + 1)
    alu = np.zeros(iwk + 2, np.float64 if double else np.float32)
    jlu = np.zeros(iwk + 2, np.int32)
    ju = np.zeros(n + 1, np.int32)
    iperm = np.zeros(2 * n, np.int32)
    ja2 = np.array(ja, np.int32)
    f = lib().orc_ilut_f64 if double else lib().orc_ilut_f32
    ierr = f(n, _i(ia), _i(ja2), _d(np.ascontiguousarray(a, np.float64)), int(lfil), C.c_double(droptol),
             C.c_double(permtol), int(pivot), int(iwk), alu.ctypes.data_as(_dp if double else _fp),
             _i(jlu), _i(ju), _i(iperm))
    return alu, jlu, ju, iperm, ja2, ierr


def lusol(alu, jlu, ju, y):
    n = ju.shape[0] - 1
    x = np.zeros(n)
    y = np.ascontiguousarray(y, np.float64)
    if alu.dtype == np.float64:
        lib().orc_lusol_f64(n, _d(y), _d(x), _d(alu), _i(jlu), _i(ju))
    else:
        lib().orc_lusol_f32(n, _d(y), _d(x), alu.ctypes.data_as(_fp), _i(jlu), _i(ju))
    return x


def rcmk(ia, ja, a=None):
    n = ia.shape[0] - 1
    nnz = int(ia[n])
    a = np.zeros(nnz) if a is None else np.ascontiguousarray(a, np.float64)
    perm = np.zeros(n, np.int32)
    rperm = np.zeros(n, np.int32)
    ib = np.zeros(n + 1, np.int32)
    jb = np.zeros(max(nnz, 1), np.int32)
    b = np.zeros(max(nnz, 1))
    nlev = lib().orc_rcmk(n, _i(ia), _i(ja), _d(a), _i(perm), _i(rperm), _i(ib), _i(jb), _d(b))
    return perm, rperm, ib, jb[:nnz], b[:nnz], nlev


def solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30, maxits=100000,
          renumbering=1, scaling=0, ilu_double=False, rtol=1e-10, precond_position=2,
          nb_fill=20, droptol=0.0, permtol=0.05):
    n = ia.shape[0] - 1
    a = np.array(a, np.float64)
    b = np.array(b, np.float64)
    x = np.zeros(n)
    p = Params(algorithm, preconditioner, precond_position, krylov, maxits, renumbering,
               scaling, int(ilu_double), rtol, nb_fill, droptol, permtol)
    info = Info()
    r = lib().orc_solve(n, _i(ia), _i(ja), _d(a), _d(b), _d(x), C.byref(p), C.byref(info))
    return x, info.as_dict(), r


# ---- reference kernels (oracle/_ref) ----------------------------------------

def to1(ia, ja):
    return (np.ascontiguousarray(ia + 1, np.int32), np.ascontiguousarray(ja + 1, np.int32))


def ref_solve(ia, ja, a, b, algorithm=8, preconditioner=6, krylov=30, maxits=100000,
              renumbering=1, scaling=0, ilu_double=False, rtol=1e-10, precond_position=2,
              nb_fill=20, droptol=0.0, permtol=0.05):
    R = ref(ilu_double)
    if R is None:
        return None
    n = ia.shape[0] - 1
    ia1, ja1 = to1(ia, ja)
    a = np.array(a, np.float64)
    b = np.array(b, np.float64)
    x = np.zeros(n)
    p = RefParams(algorithm, preconditioner, precond_position, krylov, maxits, renumbering,
                  scaling, rtol, nb_fill, droptol, permtol)
    info = RefInfo()
    r = R.ref_solve(n, _d(a), _i(ja1), _i(ia1), _d(b), _d(x), C.byref(p), C.byref(info))
    return x, info.as_dict(), r


def ref_ilu0(ia, ja, a, double=False):
    R = ref(double)
    if R is None:
        return None
    n = ia.shape[0] - 1
    ia1, ja1 = to1(ia, ja)
    nz = int(ia[n]) + n + 1
    alu = np.zeros(nz, np.float64 if double else np.float32)
    jlu = np.zeros(nz, np.int32)
    ju = np.zeros(n + 1, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    ierr = R.ref_ilu0(n, _d(a), _i(ja1), _i(ia1), alu.ctypes.data_as(_dp if double else _fp),
                      _i(jlu), _i(ju))
    return alu, jlu, ju, ierr


def ref_ilut(ia, ja, a, lfil=20, droptol=0.0, permtol=0.05, pivot=False, double=False, iwk=None):
    """The reference's ilut / ilutp, converted to the oracle's 0-based MSR."""
    R = ref(double)
    if R is None:
        return None
    n = ia.shape[0] - 1
    iwk = iwk or 2 * (n + 1) * (lfil + 1)
    ia1, ja1 = to1(ia, ja)
    alu = np.zeros(iwk + 2, np.float64 if double else np.float32)
    jlu = np.zeros(iwk + 2, np.int32)
    ju = np.zeros(n + 1, np.int32)
    iperm = np.zeros(2 * n, np.int32)
    ierr = R.ref_ilut(n, _d(np.ascontiguousarray(a, np.float64)), _i(ja1), _i(ia1), int(lfil),
                      C.c_double(droptol), C.c_double(permtol), int(pivot), int(iwk),
                      alu.ctypes.data_as(_dp if double else _fp), _i(jlu), _i(ju), _i(iperm))
    used = int(jlu[n]) - 1 if ierr == 0 else 0
    jlu0 = jlu.copy()
    jlu0[:max(used, 0)] -= 1
    return alu, jlu0, ju - 1, iperm - 1, ja1 - 1, ierr


def ref_lusol(alu, jlu, ju, y):
    R = ref(alu.dtype == np.float64)
    n = ju.shape[0] - 1
    x = np.zeros(n)
    y = np.ascontiguousarray(y, np.float64)
    R.ref_lusol(n, _d(y), _d(x), alu.ctypes.data_as(_dp if alu.dtype == np.float64 else _fp),
                _i(jlu), _i(ju))
    return x


def ref_rcmk(ia, ja, a=None):
    R = ref()
    if R is None:
        return None
    n = ia.shape[0] - 1
    nnz = int(ia[n])
    ia1, ja1 = to1(ia, ja)
    a = np.zeros(nnz) if a is None else np.ascontiguousarray(a, np.float64)
    ao = np.zeros(max(nnz, 1))
    jo = np.zeros(max(nnz, 1), np.int32)
    io = np.zeros(n + 1, np.int32)
    perm = np.zeros(n, np.int32)
    rperm = np.zeros(n, np.int32)
    nlev = R.ref_rcmk(n, _d(a), _i(ja1), _i(ia1), _d(ao), _i(jo), _i(io), _i(perm), _i(rperm))
    return perm - 1, rperm - 1, io - 1, jo[:nnz] - 1, ao[:nnz], nlev


def ref_amux(ia, ja, a, x):
    R = ref()
    n = ia.shape[0] - 1
    ia1, ja1 = to1(ia, ja)
    y = np.zeros(n)
    R.ref_amux(n, _d(np.ascontiguousarray(x, np.float64)), _d(y), _d(np.ascontiguousarray(a)),
               _i(ja1), _i(ia1))
    return y


# ---- assembly ---------------------------------------------------------------

class _Keep(list):
    def k(self, a):
        self.append(a)
        return a


def _describe(problem):
    keep = _Keep()
    p = problem
    xyz = keep.k(np.ascontiguousarray(p.xyz, np.float64))
    etype = keep.k(np.ascontiguousarray(p.etype, np.int32))
    region = keep.k(np.ascontiguousarray(p.region, np.int32))
    conn = keep.k(np.ascontiguousarray(p.conn, np.int32))
    m = OMesh(p.nnodes, _d(xyz), p.nelem, _i(etype), _i(region), _i(conn), None)
    if p.edges is not None:
        m.edges = _i(keep.k(np.ascontiguousarray(p.edges, np.int32)))
    sps = (OSpace * len(p.spaces))()
    for k, s in enumerate(p.spaces):
        sup = keep.k(np.ascontiguousarray(s.support, np.int32))
        fe = keep.k(np.ascontiguousarray(s.fixed_entity, np.int32))
        fv = keep.k(np.ascontiguousarray(s.fixed_value, np.float64))
        sps[k] = OSpace(s.bf, s.numtype, sup.shape[0], _i(sup), fe.shape[0], _i(fe), _d(fv))
    nreg = int(p.region.max()) + 1
    ts = (OTerm * len(p.terms))()
    for k, t in enumerate(p.terms):
        regs = keep.k(np.ascontiguousarray(t.regions, np.int32))
        coef = np.zeros((max(nreg, t.coef.shape[0]), t.ncoef))
        coef[: t.coef.shape[0]] = t.coef
        coef = keep.k(np.ascontiguousarray(coef))
        ts[k] = OTerm(t.kind, t.equ, t.dof, regs.shape[0], _i(regs), t.npoints, t.jac_sur, t.dim,
                      t.ncoef, coef.shape[0], _d(coef))
    return keep, m, sps, ts


def number_dofs(problem):
    keep, m, sps, ts = _describe(problem)
    nmax = problem.nmax_entity()
    tab = np.zeros((len(problem.spaces), nmax + 1), np.int32)
    N = lib().orc_number_dofs(C.byref(m), len(sps), sps, len(ts), ts, nmax, _i(tab))
    return N, tab


def assemble(problem, max_nnz=None):
    """Oracle Generate[A]: returns (ia, ja, a, b, dof_table)."""
    keep, m, sps, ts = _describe(problem)
    nmax = problem.nmax_entity()
    tab = np.zeros((len(problem.spaces), nmax + 1), np.int32)
    N = lib().orc_number_dofs(C.byref(m), len(sps), sps, len(ts), ts, nmax, _i(tab))
    if N < 0:
        raise RuntimeError("orc_number_dofs failed")
    if max_nnz is None:
        max_nnz = int(problem.nelem) * 144 + 64
    ia = np.zeros(N + 1, np.int32)
    ja = np.zeros(max_nnz, np.int32)
    a = np.zeros(max_nnz)
    b = np.zeros(N)
    nnz = lib().orc_assemble(C.byref(m), len(sps), sps, len(ts), ts, nmax, _i(tab), N, _i(ia),
                             _i(ja), _d(a), max_nnz, _d(b))
    if nnz < 0 or nnz > max_nnz:
        raise RuntimeError(f"orc_assemble failed ({nnz})")
    return ia, ja[:nnz].copy(), a[:nnz].copy(), b, tab


def element_matrix(problem, term_index, e):
    keep, m, sps, ts = _describe(problem)
    Ek = np.zeros(144)
    k = np.zeros(2, np.int32)
    r = lib().orc_element_matrix(C.byref(m), sps, C.byref(ts[term_index]), int(e), _d(Ek), _i(k))
    if r:
        raise RuntimeError(f"orc_element_matrix failed {r}")
    return Ek[: k[0] * k[1]].reshape(k[0], k[1])


def number_edges(etype, conn, visit):
    etype = np.ascontiguousarray(etype, np.int32)
    conn = np.ascontiguousarray(conn, np.int32)
    visit = np.ascontiguousarray(visit, np.int32)
    ne = etype.shape[0]
    edges = np.zeros((ne, 6), np.int32)
    mx = 6 * ne + 16
    en = np.zeros((mx, 2), np.int32)
    n = lib().orc_number_edges(ne, _i(etype), _i(conn), visit.shape[0], _i(visit), _i(edges),
                               _i(en), mx)
    return edges, en[:n]


class OPost(C.Structure):
    _fields_ = [("kind", C.c_int), ("space", C.c_int), ("nregions", C.c_int),
                ("regions", C.POINTER(C.c_int)), ("ncoef", C.c_int), ("coef_nreg", C.c_int),
                ("coef", C.POINTER(C.c_double)), ("npoints", C.c_int)]


def post(problem, x, kind, regions, space=0, coef=None, npoints=1):
    """Oracle post-processing of the solution x (original numbering): values
    (nsel, 4, 3) for local quantities, (nsel,) + total for the integral."""
    keep, m, sps, ts = _describe(problem)
    nmax = problem.nmax_entity()
    tab = np.zeros((len(problem.spaces), nmax + 1), np.int32)
    lib().orc_number_dofs(C.byref(m), len(sps), sps, len(ts), ts, nmax, _i(tab))
    regs = np.ascontiguousarray(regions, np.int32)
    cf = None if coef is None else np.ascontiguousarray(np.atleast_2d(np.asarray(coef, np.float64)))
    q = OPost(kind, space, regs.shape[0], _i(regs), 0 if cf is None else cf.shape[1],
              0 if cf is None else cf.shape[0], None if cf is None else _d(cf), npoints)
    x = np.ascontiguousarray(x, np.float64)
    dn = np.ascontiguousarray(tab[space])
    ns = lib().orc_post(C.byref(m), sps, C.byref(q), _i(dn), _d(x), None, None)
    per = 1 if kind == 4 else 12
    out = np.zeros(max(1, per * ns))
    tot = np.zeros(1)
    lib().orc_post(C.byref(m), sps, C.byref(q), _i(dn), _d(x), _d(out), _d(tot))
    vals = out[: per * ns]
    if per == 12:
        vals = vals.reshape(ns, 4, 3)
    return {"values": vals, "total": float(tot[0]), "nsel": ns}
