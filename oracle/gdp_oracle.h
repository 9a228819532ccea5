/*
 * gdp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of the reference's FP64 hot path:
 * GetDP's Galerkin assembly (src/kernel/Cal_GalerkinTermOfFemEquation.cpp,
 * Cal_AssembleTerm.cpp, DofData.cpp, Get_Geometry.cpp, src/functions/BF_*.cpp)
 * and the Sparskit solve behind LinAlg_Solve (contrib/Sparskit/Sparskit.cpp,
 * iters.f, ilut.F, matvec.f, cmkreord.f, reordering.f, unary.f, blassm.f).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library, and only as the checker.  The product never calls it.
 *
 * Conventions: CSR arrays are 0-based (ia[0] == 0).  MSR ILU storage is the
 * reference's layout shifted to 0-based indices: alu[0..n-1] hold inverted
 * diagonals, jlu[0..n] row pointers into alu/jlu (jlu[0] == n+1), ju[i] the
 * first U entry of row i, jlu[k>n] column indices.
 */
#ifndef GDP_ORACLE_H
#define GDP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- Sparskit kernels --------------------------------------------------- */
double orc_ddot(int n, const double *x, const double *y);
void orc_set_dot_blocks(int nblocks);
void orc_set_threads(int nthreads); /* CPU baselines: OpenMP threads (1: sequential) */
void orc_amux(int n, const double *x, double *y, const int *ia, const int *ja,
              const double *a);
int orc_ilu0_f32(int n, const int *ia, const int *ja, const double *a,
                 float *alu, int *jlu, int *ju);
/* ilut (ilut.F:52-396, pivot 0) / ilutp (:397-806, pivot 1, mbloc = n):
 * threshold ILU with lfil entries per L/U row, dropping below droptol
 * (relative to the row norm for U), column pivoting with permtol.  0-based
 * MSR output (alu[0..n) inverted diagonal, jlu[0..n] row starts, ju row
 * starts of U) with room for iwk entries; ilutp permutes ja in place and
 * fills iperm (2n).  Returns 0 or Sparskit's ierr (-1..-5). */
int orc_ilut_f32(int n, const int *ia, int *ja, const double *a, int lfil, double droptol,
                 double permtol, int pivot, int iwk, float *alu, int *jlu, int *ju, int *iperm);
int orc_ilut_f64(int n, const int *ia, int *ja, const double *a, int lfil, double droptol,
                 double permtol, int pivot, int iwk, double *alu, int *jlu, int *ju, int *iperm);
int orc_ilu0_f64(int n, const int *ia, const int *ja, const double *a,
                 double *alu, int *jlu, int *ju);
void orc_lusol_f32(int n, const double *y, double *x, const float *alu,
                   const int *jlu, const int *ju);
void orc_lusol_f64(int n, const double *y, double *x, const double *alu,
                   const int *jlu, const int *ju);
/* RCMK: fills perm[old] = new and rperm[new] = old (0-based) and the permuted
 * matrix with sorted columns.  Returns the number of BFS levels. */
int orc_rcmk(int n, const int *ia, const int *ja, const double *a, int *perm,
             int *rperm, int *ia_out, int *ja_out, double *a_out);

/* solver.par subset (contrib/Sparskit/Sparskit.cpp:1847-1868). */
typedef struct {
  int algorithm;        /* 1 CG, 8 GMRES, 9 FGMRES */
  int preconditioner;   /* 0 NONE, 1 ILUT, 2 ILUTP, 6 ILU0, 8 DIAGONAL */
  int precond_position; /* 1 left, 2 right (default 2) */
  int krylov;           /* Krylov_Size */
  int maxits;           /* Nb_Iter_Max */
  int renumbering;      /* 0 none, 1 RCMK */
  int scaling;          /* 0 none, 1 diagonal */
  int ilu_double;       /* 0: FP32 factor storage (reference default) */
  double rtol;          /* Stopping_Test */
  int nb_fill;          /* ILUT/ILUTP: Nb_Fill */
  double droptol;       /* Dropping_Tolerance */
  double permtol;       /* ILUTP: Permutation_Tolerance */
} orc_params;

typedef struct {
  int its;         /* ipar(7) at exit */
  int status;      /* ipar(1) at exit: 0 converged, -1 max its, -3 breakdown */
  int ilu_ierr;    /* ILU(0) zero-pivot row (1-based) or 0 */
  double res_est;  /* fpar(6): estimated residual norm at exit */
  double res0;     /* norm of the initial residual (= |b| after permutation) */
  double true_res; /* |A sol - rhs|_2 of the solved (permuted, scaled) system */
} orc_info;

/* Full solve_matrix flow.  a is modified in place when scaling is on, as in
 * the reference.  Returns 0, or the ILU zero-pivot row, or -100 (zero row). */
int orc_solve(int n, const int *ia, const int *ja, double *a, double *b,
              double *x, const orc_params *p, orc_info *info);

/* Plain Krylov drivers on an explicit operator, used to pin the restatement
 * against the reference kernels on arbitrary inputs. */
typedef void (*orc_op)(void *ctx, const double *in, double *out);
int orc_gmres(int n, const double *rhs, double *sol, int m, int maxits,
              double rtol, orc_op matvec, orc_op precond, void *ctx,
              orc_info *info);
int orc_fgmres(int n, const double *rhs, double *sol, int m, int maxits,
               double rtol, orc_op matvec, orc_op precond, void *ctx,
               orc_info *info);
int orc_cg(int n, const double *rhs, double *sol, int maxits, double rtol,
           orc_op matvec, orc_op precond, void *ctx, orc_info *info);
int orc_bcgstab(int n, const double *rhs, double *sol, int maxits, double rtol,
                orc_op matvec, orc_op precond, void *ctx, orc_info *info);
int orc_tfqmr(int n, const double *rhs, double *sol, int maxits, double rtol,
              orc_op matvec, orc_op precond, void *ctx, orc_info *info);

/* ---- Mesh entities and DOF numbering ------------------------------------ */
/* Element types use Gmsh codes: 1 line, 2 triangle, 4 tetrahedron. */
enum { ORC_LINE = 1, ORC_TRI = 2, ORC_TET = 4 };

/* Signed edge numbers per element (GeoEntity.cpp:437-490): elements are
 * visited in the order given by `visit` (indices into the element arrays);
 * each new edge (sorted global node pair) gets the next number; the sign is
 * +1 when the local edge runs from the lower to the higher node number.
 * edges_out[e*6 + k] receives the signed number of local edge k of element e
 * (0 for elements never visited).  edge_nodes receives (lo, hi) per edge.
 * Returns the number of edges. */
int orc_number_edges(int ne, const int *etype, const int *conn, int nvisit,
                     const int *visit, int *edges_out, int *edge_nodes,
                     int max_edges);

/* ---- Assembly ----------------------------------------------------------- */
enum {
  ORC_BF_NODE = 0, /* BF_Node on NodesOf */
  ORC_BF_EDGE = 1, /* BF_Edge on EdgesOf */
  ORC_BF_PERP_EDGE = 2 /* BF_PerpendicularEdge on NodesOf (2D, Form1P) */
};
enum {
  /* bilinear terms (Equ space, Dof space) */
  ORC_T_GRAD_GRAD = 1,   /* [ coef * Dof{d u}, {d v} ]  FORM1 x FORM1     */
  ORC_T_CURL_CURL = 2,   /* [ coef * Dof{d a}, {d a} ]  FORM2 (3D edges)  */
  ORC_T_EDGE_MASS = 3,   /* [ coef * Dof{a}, {a} ]      FORM1 (3D edges)  */
  ORC_T_NODE_MASS = 4,   /* [ coef * Dof{u}, {u} ]      FORM0             */
  /* source terms without Dof{} (Dof_InitDofForNoDof: fixed pseudo-DOF 1) */
  ORC_T_SRC_NODE = 10,   /* [ f, {u} ]  scalar f, FORM0 test              */
  ORC_T_SRC_EDGE = 11    /* [ f, {a} ]  vector f, FORM1 test              */
};
/* On an ORC_BF_PERP_EDGE space the edge kinds take the 2D perpendicular forms:
 * CURL_CURL is FORM2P x FORM2P (ChangeOfCoord_Form2P, Cal_Product123),
 * EDGE_MASS and SRC_EDGE are FORM1P (ChangeOfCoord_Form1P, Cal_Product3),
 * Get_Geometry.cpp:1929-1947, 1965-1976. */

typedef struct {
  int bf;           /* ORC_BF_NODE / ORC_BF_EDGE */
  int numtype;      /* BasisFunction.Num: global DOF key, parse order */
  int nsupport;     /* regions of the support */
  const int *support;
  int nfixed;       /* constrained entities (node numbers 1-based / edge numbers) */
  const int *fixed_entity;
  const double *fixed_value;
} orc_space;

typedef struct {
  int kind;
  int equ, dof;      /* space indices (dof = -1 for source terms) */
  int nregions;      /* "In" regions */
  const int *regions;
  int npoints;       /* Gauss points: tri 1/3/4, tet 1/4 */
  int jac_sur;       /* 1: Jacobian Sur (triangles embedded in 3D) */
  int dim;           /* 2 or 3: dimension of the vector forms */
  int ncoef;         /* number of values per region: 1 (scalar), 3, 9 (tensor) */
  int coef_nreg;     /* coefficient table covers region ids 0..coef_nreg-1 */
  const double *coef;/* coef[r*ncoef + c]: the whole coefficient, sign included */
} orc_term;

typedef struct {
  int nnodes;
  const double *xyz;     /* 3 per node */
  int nelem;
  const int *etype;
  const int *region;
  const int *conn;       /* 4 per element (node numbers, 1-based) */
  const int *edges;      /* 6 per element (signed edge numbers) or NULL */
} orc_mesh;

/* DOF numbering (DofData.cpp:36-45, 1279-1323): every (space, entity) that
 * appears in an element of some term's regions within the space's support is
 * a DOF, fixed when constrained; unknowns are numbered 1..N in (numtype,
 * entity) order.  dofnum_out[s][entity] is filled for each space s:
 * > 0 unknown number, < 0 -(fixed index+1), 0 absent.  Returns N. */
int orc_number_dofs(const orc_mesh *m, int nspaces, const orc_space *sp,
                    int nterms, const orc_term *t, int nentities_max,
                    int *dofnum_out);

/* Generate[A] (SolvingOperations.cpp:276 -> Treatment_FemFormulation):
 * element-major loop, per element every term in order, Ek by Gauss
 * quadrature, scatter to A/b.  dofnum is the table from orc_number_dofs.
 * Returns nnz; ia (N+1), ja/a (nnz, up to max_nnz) and b (N) are written. */
long orc_assemble(const orc_mesh *m, int nspaces, const orc_space *sp,
                  int nterms, const orc_term *t, int nentities_max,
                  const int *dofnum, int N, int *ia, int *ja, double *a,
                  long max_nnz, double *b);

/* Post-processing (Pos_LocalOrIntegralQuantity, Cal_PostQuantity.cpp:29-247,
 * interpolation as Pos_FemInterpolation.cpp:226-330) of the solution x on the
 * elements of q->regions, in element order.  dofnum: the table of q->space
 * (orc_number_dofs).  out: 12 values per element (vertex-major, 3 components)
 * for VALUE, D, COEF_D; one per element for INT_COEF_D2, *total = their sum in
 * element order (Cal_PostCumulativeQuantity).  Returns the element count. */
enum { ORC_PQ_VALUE = 1, ORC_PQ_D = 2, ORC_PQ_COEF_D = 3, ORC_PQ_INT_COEF_D2 = 4 };
typedef struct {
  int kind, space;
  int nregions;
  const int *regions;
  int ncoef, coef_nreg;
  const double *coef;
  int npoints;
} orc_post_q;
int orc_post(const orc_mesh *m, const orc_space *sp, const orc_post_q *q, const int *dofnum,
             const double *x, double *out, double *total);

/* Element matrix of one term on one element (for closed-form tests). */
int orc_element_matrix(const orc_mesh *m, const orc_space *sp, const orc_term *t, int e,
                       double *Ek /* k*k row-major */, int *k_out);

#ifdef __cplusplus
}
#endif
#endif
