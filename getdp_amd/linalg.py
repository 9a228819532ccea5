"""Python mirror of the reference's LinAlg interface (src/kernel/LinAlg.h:90-213)
over the device backend's C-ABI.

Names follow the reference: ``LinAlg_CreateSolver`` -> :class:`Solver`,
``LinAlg_CreateVector`` -> :class:`Vector`, ``LinAlg_CreateMatrix`` ->
:class:`Matrix`, ``LinAlg_Solve`` -> :func:`solve`, ``LinAlg_SolveAgain`` ->
:func:`solve_again`, ``LinAlg_ProdMatrixVector`` -> :func:`prod_matrix_vector`;
``Generate[A]`` of a formulation -> :meth:`FESystem.generate`.  Errors raise
:class:`GdpError` carrying the backend's status code and message.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from ._lib import (GdpError, MeshDesc, PostDesc, SolveInfo, SpaceDesc, TermDesc, check, dptr, iptr,
                   lib)

__all__ = ["init", "Solver", "Vector", "Matrix", "FESystem", "solve", "solve_again",
           "prod_matrix_vector", "GdpError"]


def init(device: int = 0):
    check(lib().gdp_init(device), "gdp_init")


class _Handle:
    _destroy = None

    def __init__(self):
        self.h = C.c_void_p()

    def close(self):
        if self.h and self.h.value:
            getattr(lib(), self._destroy)(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver(_Handle):
    """Solver_Params (+ cached renumbering/factorisation), LinAlg_CreateSolver."""
    _destroy = "gdp_solver_destroy"

    def __init__(self, solver_par: Optional[str] = None, **options):
        super().__init__()
        path = solver_par.encode() if solver_par else None
        check(lib().gdp_solver_create(path, C.byref(self.h)), "gdp_solver_create")
        for k, v in options.items():
            self.set(k, v)

    def set(self, name: str, value):
        check(lib().gdp_solver_set(self.h, name.encode(), str(value).encode()), f"set {name}")

    def get(self, name: str) -> float:
        v = C.c_double()
        check(lib().gdp_solver_get(self.h, name.encode(), C.byref(v)), f"get {name}")
        return v.value

    def info(self) -> dict:
        s = SolveInfo()
        check(lib().gdp_solver_info(self.h, C.byref(s)), "gdp_solver_info")
        return s.as_dict()


class Vector(_Handle):
    _destroy = "gdp_vector_destroy"

    def __init__(self, n: int, values: Optional[np.ndarray] = None):
        super().__init__()
        self.n = int(n)
        check(lib().gdp_vector_create(self.n, C.byref(self.h)), "gdp_vector_create")
        if values is not None:
            self.set(values)

    def set(self, values):
        v = np.ascontiguousarray(values, np.float64)
        check(lib().gdp_vector_set(self.h, dptr(v), v.shape[0]), "gdp_vector_set")

    def get(self) -> np.ndarray:
        out = np.zeros(self.n)
        check(lib().gdp_vector_get(self.h, dptr(out), self.n), "gdp_vector_get")
        return out

    def zero(self):
        check(lib().gdp_vector_zero(self.h), "gdp_vector_zero")

    def add_double(self, i: int, d: float):
        check(lib().gdp_vector_add_double(self.h, int(i), float(d)), "gdp_vector_add_double")

    def assemble(self):
        check(lib().gdp_vector_assemble(self.h), "gdp_vector_assemble")

    def norm2(self) -> float:
        v = C.c_double()
        check(lib().gdp_vector_norm2(self.h, C.byref(v)), "gdp_vector_norm2")
        return v.value


class Matrix(_Handle):
    _destroy = "gdp_matrix_destroy"

    def __init__(self, n: int):
        super().__init__()
        self.n = int(n)
        check(lib().gdp_matrix_create(self.n, self.n, C.byref(self.h)), "gdp_matrix_create")

    def zero(self):
        check(lib().gdp_matrix_zero(self.h), "gdp_matrix_zero")

    def add_double(self, i: int, j: int, d: float):
        check(lib().gdp_matrix_add_double(self.h, int(i), int(j), float(d)), "add_double")

    def add_coo(self, i, j, d):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        d = np.ascontiguousarray(d, np.float64)
        check(lib().gdp_matrix_add_coo(self.h, i.shape[0], iptr(i), iptr(j), dptr(d)), "add_coo")

    def assemble(self):
        check(lib().gdp_matrix_assemble(self.h), "gdp_matrix_assemble")

    def set_csr(self, ia, ja, a):
        ia = np.ascontiguousarray(ia, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        a = np.ascontiguousarray(a, np.float64)
        check(lib().gdp_matrix_set_csr(self.h, self.n, iptr(ia), iptr(ja), dptr(a)), "set_csr")

    @property
    def nnz(self) -> int:
        v = C.c_long()
        check(lib().gdp_matrix_nnz(self.h, C.byref(v)), "gdp_matrix_nnz")
        return v.value

    def get_csr(self):
        nnz = self.nnz
        ia = np.zeros(self.n + 1, np.int32)
        ja = np.zeros(max(nnz, 1), np.int32)
        a = np.zeros(max(nnz, 1))
        check(lib().gdp_matrix_get_csr(self.h, iptr(ia), iptr(ja), dptr(a)), "get_csr")
        return ia, ja[:nnz], a[:nnz]


def prod_matrix_vector(A: Matrix, x: Vector, y: Vector):
    check(lib().gdp_prod_matrix_vector(A.h, x.h, y.h), "gdp_prod_matrix_vector")


def sell_info(A: Matrix) -> dict:
    """The matrix's SpMV layout: padded entries and whether the columns are
    stored as 16-bit offsets."""
    npad = C.c_long()
    comp = C.c_int()
    check(lib().gdp_matrix_sell_info(A.h, C.byref(npad), C.byref(comp)), "gdp_matrix_sell_info")
    return {"nnz_pad": npad.value, "compressed": bool(comp.value)}


def bench_solver_spmv(s: Solver, reps: int = 20) -> dict:
    """Average device time (ms) of the SpMV on the solver's renumbered copy of
    the last matrix it solved (the operator of its Krylov iteration)."""
    ms = C.c_double()
    comp = C.c_int()
    check(lib().gdp_solver_bench_spmv(s.h, int(reps), C.byref(ms), C.byref(comp)), "gdp_solver_bench_spmv")
    return {"ms": ms.value, "compressed": bool(comp.value)}


def solver_prec_apply(s: Solver, v: Vector, out: Vector, reps: int = 1) -> float:
    """out = (LU)^{-1} v with the ILU(0) factors of the solver's last solve, in
    its internal (renumbered, scaled) numbering; returns ms per application."""
    ms = C.c_double()
    check(lib().gdp_solver_prec_apply(s.h, v.h, out.h, int(reps), C.byref(ms)), "gdp_solver_prec_apply")
    return ms.value


def bench_spmv(A: Matrix, x: Vector, y: Vector, reps: int = 20) -> float:
    """Average device time (ms) of one CSR SpMV launch over `reps` launches."""
    v = C.c_double()
    check(lib().gdp_bench_spmv(A.h, x.h, y.h, int(reps), C.byref(v)), "gdp_bench_spmv")
    return v.value


def solve(A: Matrix, b: Vector, s: Solver, x: Vector) -> dict:
    check(lib().gdp_solve(A.h, b.h, s.h, x.h), "gdp_solve")
    return s.info()


def solve_again(A: Matrix, b: Vector, s: Solver, x: Vector) -> dict:
    check(lib().gdp_solve_again(A.h, b.h, s.h, x.h), "gdp_solve_again")
    return s.info()


def res_write(path: str, x: Vector, create: bool = True, binary: bool = False, num: int = 0,
              time: float = 0.0, time_imag: float = 0.0, step: int = 0):
    """SaveSolution: one $Solution block (Dof_WriteFileRES, DofData.cpp:741-767),
    after the $ResFormat block of a new file when `create`."""
    check(lib().gdp_res_write(path.encode(), int(create), int(binary), num, time, time_imag, step, x.h),
          "gdp_res_write")


def res_read(path: str, x: Vector, num: int = 0, which: int = -1) -> dict:
    """Solution `which` (the last when < 0) of DofData `num` from a .res file
    (Dof_ReadFileRES, DofData.cpp:922-1010) into x."""
    t, ti = C.c_double(), C.c_double()
    st, cnt = C.c_int(), C.c_int()
    check(lib().gdp_res_read(path.encode(), num, which, x.h, C.byref(t), C.byref(ti), C.byref(st),
                             C.byref(cnt)), "gdp_res_read")
    return {"time": t.value, "time_imag": ti.value, "step": st.value, "count": cnt.value}


# post-processing quantity kinds (gdp_hip.h GDP_PQ_*)
PQ_VALUE, PQ_D, PQ_COEF_D, PQ_INT_COEF_D2 = 1, 2, 3, 4


class FESystem(_Handle):
    """One formulation's device assembly (PRE pass at construction,
    ``generate`` = Generate[A])."""
    _destroy = "gdp_fe_destroy"

    def __init__(self, problem):
        super().__init__()
        p = problem
        self._keep = []

        def keep(a):
            self._keep.append(a)
            return a

        xyz = keep(np.ascontiguousarray(p.xyz, np.float64))
        etype = keep(np.ascontiguousarray(p.etype, np.int32))
        region = keep(np.ascontiguousarray(p.region, np.int32))
        conn = keep(np.ascontiguousarray(p.conn, np.int32))
        md = MeshDesc(p.nnodes, dptr(xyz), p.nelem, iptr(etype), iptr(region), iptr(conn), None, 0)
        if p.edges is not None:
            edges = keep(np.ascontiguousarray(p.edges, np.int32))
            md.edges = iptr(edges)
            md.nedges = int(p.edge_nodes.shape[0])
        sds = (SpaceDesc * len(p.spaces))()
        for k, s in enumerate(p.spaces):
            sup = keep(np.ascontiguousarray(s.support, np.int32))
            fe = keep(np.ascontiguousarray(s.fixed_entity, np.int32))
            fv = keep(np.ascontiguousarray(s.fixed_value, np.float64))
            sds[k] = SpaceDesc(s.bf, s.numtype, sup.shape[0], iptr(sup), fe.shape[0], iptr(fe),
                               dptr(fv))
        nreg = int(p.region.max()) + 1
        tds = (TermDesc * len(p.terms))()
        for k, t in enumerate(p.terms):
            regs = keep(np.ascontiguousarray(t.regions, np.int32))
            coef = np.zeros((max(nreg, t.coef.shape[0]), t.ncoef))
            coef[: t.coef.shape[0]] = t.coef
            coef = keep(np.ascontiguousarray(coef))
            tds[k] = TermDesc(t.kind, t.equ, t.dof, regs.shape[0], iptr(regs), t.npoints,
                              t.jac_sur, t.ncoef, coef.shape[0], dptr(coef))
        self._mesh = md
        check(lib().gdp_fe_create(C.byref(md), len(p.spaces), sds, len(p.terms), tds,
                                  C.byref(self.h)), "gdp_fe_create")
        self._keep = []
        self.nmax = max(p.nnodes, 0 if p.edge_nodes is None else int(p.edge_nodes.shape[0]))
        n = C.c_int()
        nnz = C.c_long()
        check(lib().gdp_fe_num_dofs(self.h, C.byref(n), C.byref(nnz)), "gdp_fe_num_dofs")
        self.N, self.nnz = n.value, nnz.value
        self.nspaces = len(p.spaces)

    def dof_table(self, space: int = 0) -> np.ndarray:
        out = np.zeros(self.nmax + 1, np.int32)
        check(lib().gdp_fe_dof_table(self.h, space, iptr(out), self.nmax), "gdp_fe_dof_table")
        return out

    def write_pre(self, path: str, num: int = 0, resolution_index: int = 0, system_index: int = 0,
                  fs_index=None):
        """The $DofData block of a .pre file (Dof_WriteFilePRE) for these DOFs."""
        libc = C.CDLL(None)
        libc.fopen.restype = C.c_void_p
        libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        libc.fclose.argtypes = [C.c_void_p]
        fs = np.ascontiguousarray(fs_index if fs_index is not None else np.arange(self.nspaces),
                                  np.int32)
        f = libc.fopen(path.encode(), b"w")
        if not f:
            raise OSError(f"cannot open {path}")
        try:
            check(lib().gdp_fe_write_pre(self.h, f, num, resolution_index, system_index, fs.shape[0],
                                         fs.ctypes.data_as(C.POINTER(C.c_int))), "gdp_fe_write_pre")
        finally:
            libc.fclose(f)

    def pattern(self):
        ia = np.zeros(self.N + 1, np.int32)
        ja = np.zeros(max(self.nnz, 1), np.int32)
        check(lib().gdp_fe_pattern(self.h, iptr(ia), iptr(ja)), "gdp_fe_pattern")
        return ia, ja[: self.nnz]

    def generate(self, A: Matrix, b: Vector):
        check(lib().gdp_fe_generate(self.h, A.h, b.h), "gdp_fe_generate")

    def timing(self) -> dict:
        t = [C.c_double() for _ in range(3)]
        check(lib().gdp_fe_last_timing(self.h, *[C.byref(x) for x in t]), "timing")
        return dict(ms_total=t[0].value, ms_geometry=t[1].value, ms_gather=t[2].value)

    def post(self, x: Vector, kind: int, regions, space: int = 0, coef=None, npoints: int = 1) -> dict:
        """PostOperation Print[...] of a local or integral quantity of the
        solution x on the elements of `regions` (element order): kind
        PQ_VALUE / PQ_D / PQ_COEF_D -> values (nsel, 4, 3); PQ_INT_COEF_D2 ->
        values (nsel,) and their total.  coef: per-region table (nreg, ncoef)."""
        regs = np.ascontiguousarray(regions, np.int32)
        cf = None if coef is None else np.ascontiguousarray(np.atleast_2d(np.asarray(coef, np.float64)))
        d = PostDesc(kind, space, regs.shape[0], iptr(regs), 0 if cf is None else cf.shape[1],
                     0 if cf is None else cf.shape[0], None if cf is None else dptr(cf), npoints)
        ns = C.c_int()
        check(lib().gdp_fe_post(self.h, x.h, C.byref(d), C.byref(ns), None, 0, None, None), "gdp_fe_post")
        per = 1 if kind == PQ_INT_COEF_D2 else 12
        out = np.zeros(max(1, per * ns.value))
        tot = C.c_double()
        ms = C.c_double()
        check(lib().gdp_fe_post(self.h, x.h, C.byref(d), C.byref(ns), dptr(out), out.shape[0],
                                C.byref(tot), C.byref(ms)), "gdp_fe_post")
        vals = out[: per * ns.value]
        if per == 12:
            vals = vals.reshape(ns.value, 4, 3)
        return {"values": vals, "total": tot.value, "kernel_ms": ms.value, "nsel": ns.value}


def assemble(problem, fe: Optional[FESystem] = None):
    """Convenience: build the FE system, generate A and b on the device and
    return (fe, A, b)."""
    fe = fe or FESystem(problem)
    A = Matrix(fe.N)
    b = Vector(fe.N)
    fe.generate(A, b)
    return fe, A, b
