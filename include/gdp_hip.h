/*
 * gdp_hip.h -- C-ABI of the MI355X (gfx950) finite-element backend.
 *
 * This library is a LinAlg backend in the sense of GetDP's src/kernel/LinAlg.h
 * (the typedef switch at LinAlg.h:20-88 selects LinAlg_SPARSKIT.cpp or
 * LinAlg_PETSC.cpp; this is a third branch), plus a batched device path for
 * the per-element Galerkin assembly of src/kernel/Cal_GalerkinTermOfFemEquation.cpp.
 *
 * All entry points are `extern "C"`, take plain pointers and sizes, and return
 * an int status: GDP_OK (0) or a negative GDP_E_* code; gdp_last_error()
 * returns the message (the reference reports the same conditions through
 * Message::Error / Message::Warning, e.g. contrib/Sparskit/Sparskit.cpp:429-453,
 * 556-567).  Indices at this interface are 0-based, as at the LinAlg interface
 * (DofData.cpp:1477 passes NumDof-1).
 *
 * Every device entry point runs on the library's single HIP stream and
 * returns after the work it enqueued has completed, unless documented
 * otherwise.  The library is not re-entrant (the reference is single-threaded,
 * LinAlg_SPARSKIT.cpp:27 keeps static solver options).
 */
#ifndef GDP_HIP_H
#define GDP_HIP_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GDP_OK 0
#define GDP_E_ARG -1        /* invalid argument                               */
#define GDP_E_HIP -2        /* HIP runtime error                              */
#define GDP_E_STATE -3      /* call out of sequence                           */
#define GDP_E_UNSUPPORTED -4/* term/element not supported by the device path  */
#define GDP_E_ZERO_PIVOT -5 /* ILU zero pivot (row in gdp_solve_info.ilu_row)  */
#define GDP_E_ZERO_ROW -6   /* "Zero row in matrix" (Sparskit.cpp:220-223)    */
#define GDP_E_NOMEM -7      /* device or host allocation failure              */
#define GDP_E_TIMEOUT -8    /* a device wait exceeded its time bound          */
#define GDP_E_NOT_CONVERGED -9 /* Krylov stopped without reaching the tolerance */

/* ---- library / device ------------------------------------------------------ */
/* Select the device and create the stream.  Replaces LinAlg_InitializeSolver
 * (LinAlg_SPARSKIT.cpp:29-67) as the first call. */
int gdp_init(int device);
int gdp_finalize(void);
const char *gdp_last_error(void);
int gdp_device_synchronize(void);

/* ---- solver parameters (Solver_Params, contrib/Sparskit/Sparskit.h:83-103) */
typedef struct gdp_solver gdp_solver;
/* LinAlg_CreateSolver (LinAlg_SPARSKIT.cpp:74-110): reads a solver.par file
 * (NULL: defaults of Sparskit.cpp:1847-1868).  Unknown file -> defaults. */
int gdp_solver_create(const char *solver_par_file, gdp_solver **s);
/* init_solver_option (Sparskit.cpp:2000-2042): set one parameter by name,
 * e.g. ("Preconditioner", "6").  Extra keys of this backend:
 *   Precision           0 FP64 (default), 1 FP32 inner solve + FP64
 *                        iterative refinement
 *   ILU_Ordering        0 as the matrix is ordered after Renumbering_Technique
 *                        (default: the reference's ordering), 1 multicolour
 *   Orthogonalization   0 modified Gram-Schmidt with selective
 *                        re-orthogonalisation (iters.f:3475, default),
 *                        1 classical Gram-Schmidt applied twice
 *   Refinement_Steps    max outer steps for Precision 1 (default 20)
 *   Inner_Tolerance     inner relative tolerance for Precision 1 (1e-6)
 *   Precision 2         FP64 flexible GMRES (residuals, Gram-Schmidt,
 *                        updates) preconditioned by Inner_Krylov_Size (20)
 *                        steps of FP32 GMRES with FP32 SpMV and the FP32
 *                        ILU(0) of Preconditioner 6; its = outer steps   */
int gdp_solver_set(gdp_solver *s, const char *name, const char *value);
int gdp_solver_get(const gdp_solver *s, const char *name, double *value);
int gdp_solver_destroy(gdp_solver *s); /* LinAlg_DestroySolver */

typedef struct {
  int its;            /* matrix-vector products, ipar(7) (iters.f)         */
  int status;         /* ipar(1) at exit: 0 converged, -1 max its, -3 ...  */
  int ilu_row;        /* zero-pivot row (1-based) or 0                     */
  int ilu_levels;     /* dependency depth of the L factor                  */
  int outer_its;      /* iterative-refinement steps (Precision 1)          */
  double res0;        /* initial residual norm                             */
  double res_est;     /* estimated residual norm at exit (fpar(6))         */
  double true_res;    /* |A x - b|_2 (solved system) recomputed in FP64     */
  double rel_res;     /* true_res / |b|_2                                  */
  double t_setup_ms;  /* renumbering + factorisation                       */
  double t_solve_ms;  /* Krylov iterations                                 */
  int ilu_sweep;      /* kernels applying the ILU factors: 0 none,
                         GDP_SWEEP_FREE sync-free multi-wave sweeps,
                         GDP_SWEEP_WAVE one-wave level walk            */
  double t_factor_ms; /* part of t_setup_ms: the factorisation (ILUT/ILUTP
                         on the host, ILU(0) on the device)            */
  double t_load_ms;   /* part of t_setup_ms: ILUT/ILUTP factors to the device,
                         sweep analysis and packing                     */
} gdp_solve_info;
enum { GDP_SWEEP_FREE = 1, GDP_SWEEP_WAVE = 2 };
int gdp_solver_info(const gdp_solver *s, gdp_solve_info *info);

/* ---- vectors (gVector) --------------------------------------------------- */
typedef struct gdp_vector gdp_vector;
int gdp_vector_create(int n, gdp_vector **v);                 /* LinAlg_CreateVector  */
int gdp_vector_destroy(gdp_vector *v);                        /* LinAlg_DestroyVector */
int gdp_vector_zero(gdp_vector *v);                           /* LinAlg_ZeroVector    */
int gdp_vector_size(const gdp_vector *v, int *n);             /* LinAlg_GetVectorSize */
int gdp_vector_set(gdp_vector *v, const double *host, int n); /* LinAlg_SetDoubleInVector, bulk */
int gdp_vector_get(const gdp_vector *v, double *host, int n); /* LinAlg_GetDoubleInVector, bulk */
/* LinAlg_AddDoubleInVector (LinAlg_SPARSKIT.cpp:356-364): staged on the host
 * and applied in call order by gdp_vector_assemble, or by the first entry
 * point that reads the vector (arithmetic, norms, dot, get, copy, products,
 * solves, output), so that every read sees the additions as the reference's
 * direct `V[i] += d` would.  Entry points that overwrite every entry
 * (gdp_vector_set, _set_all, _zero, _scan, _read, and outputs of products
 * and solves) drop pending additions, as the reference's overwrite does. */
int gdp_vector_add_double(gdp_vector *v, int i, double d);
int gdp_vector_assemble(gdp_vector *v);                       /* LinAlg_AssembleVector */
int gdp_vector_norm2(const gdp_vector *v, double *norm);      /* LinAlg_VectorNorm2   */
int gdp_vector_copy(const gdp_vector *src, gdp_vector *dst);  /* LinAlg_CopyVector    */
/* Device address of the data (FP64, n entries) for zero-copy interop. */
int gdp_vector_device_ptr(gdp_vector *v, double **dptr);

/* Whole-vector arithmetic (LinAlg_SPARSKIT.cpp over Sparskit.cpp:797-1215),
 * on the device, with the reference's argument rules (the result must alias
 * the operand named, else GDP_E_ARG "Wrong arguments in ..."):
 *   gdp_vector_add        v3 == v1: v1 += v2          (LinAlg_AddVectorVector)
 *   gdp_vector_add_scaled v3 == v1: v1 += d * v2      (LinAlg_AddVectorProdVectorDouble)
 *   gdp_vector_sub        v3 == v1: v1 -= v2; v3 == v2: v2 = v1 - v2 (LinAlg_SubVectorVector)
 *   gdp_vector_scale      v2 == v1: v1 *= d           (LinAlg_ProdVectorDouble)
 *   gdp_vector_swap, gdp_vector_set_all (every entry = d)  (LinAlg_SwapVector, LinAlg_SetVector)
 *   gdp_vector_dot        (LinAlg_ProdVectorVector; fixed two-stage reduction)
 *   gdp_vector_norm_inf   (LinAlg_VectorNormInf; NaN entries ignored, as there)
 *   gdp_vector_dummy      entries with dummy[i] == 1 set to 0 (LinAlg_DummyVector;
 *                         dummy = Current.DofData->DummyDof, NULL: none)
 * Elementwise results equal the sequential loops bit for bit. */
int gdp_vector_add(gdp_vector *v1, gdp_vector *v2, gdp_vector *v3);
int gdp_vector_add_scaled(gdp_vector *v1, gdp_vector *v2, double d, gdp_vector *v3);
int gdp_vector_sub(gdp_vector *v1, gdp_vector *v2, gdp_vector *v3);
int gdp_vector_scale(gdp_vector *v1, double d, gdp_vector *v2);
int gdp_vector_swap(gdp_vector *v1, gdp_vector *v2);
int gdp_vector_set_all(gdp_vector *v, double d);
int gdp_vector_dot(const gdp_vector *v1, const gdp_vector *v2, double *d);
/* gdp_vector_dot without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, code that runs beside graph capture).  Same kernels and
 * summation order as gdp_vector_dot, so the same bits.  Its 512 partial sums
 * live in a fixed 4 KiB device array, one per GPU, held until the process
 * ends; vectors of any length fit (the reduction always runs 512 blocks).
 * Needs gdp_init first; refuses vectors with pending gdp_vector_add_double
 * additions (GDP_E_STATE), since applying them needs device memory.  Calls
 * that overlap on one GPU share the array and race. */
int gdp_vector_dot_noalloc(const gdp_vector *v1, const gdp_vector *v2, double *d);
int gdp_vector_norm_inf(const gdp_vector *v, double *norm);
int gdp_vector_dummy(gdp_vector *v, const int *dummy);
/* Vector text and binary I/O (LinAlg_PrintVector / LinAlg_WriteVector /
 * LinAlg_ScanVector / LinAlg_ReadVector, LinAlg_SPARSKIT.cpp:174-229): print
 * writes one "%.16g\n" per entry, write the n doubles in native byte order
 * and a newline; scan reads n numbers as fscanf("%lf"), read n raw doubles. */
int gdp_vector_print(const gdp_vector *v, FILE *f);
int gdp_vector_write(const gdp_vector *v, FILE *f);
int gdp_vector_scan(gdp_vector *v, FILE *f);
int gdp_vector_read(gdp_vector *v, FILE *f);
/* Result files (.res).  gdp_res_write appends one solution block
 * (Dof_WriteFileRES, DofData.cpp:741-767: "$Solution  /" "* DofData #num *" "/",
 * "num time time_imag step", the vector, "$EndSolution"), after writing the
 * $ResFormat block to a new file when create != 0 (Dof_WriteFileRES0,
 * :631-642).  gdp_res_read follows Dof_ReadFileRES (:922-1010) for DofData
 * `num` (any when < 0): solution `which` of those found (the last when < 0)
 * goes into x; *count = solutions found.  ASCII or binary per the file's
 * $ResFormat; format version 1.0 files carry no imaginary time. */
int gdp_res_write(const char *path, int create, int binary, int num, double time, double time_imag,
                  int step, const gdp_vector *x);
int gdp_res_read(const char *path, int num, int which, gdp_vector *x, double *time, double *time_imag,
                 int *step, int *count);

/* ---- matrices (gMatrix) -------------------------------------------------- */
typedef struct gdp_matrix gdp_matrix;
/* Whole-matrix arithmetic (add_matrix_matrix, add_matrix_prod_matrix_double,
 * prod_matrix_double, Sparskit.cpp:865-1135), same argument rules:
 *   gdp_matrix_add        m3 == m1: m1 += m2; m3 == m2: m2 += m1 (LinAlg_AddMatrixMatrix)
 *   gdp_matrix_add_scaled m3 == m1: m1 += d * m2      (LinAlg_AddMatrixProdMatrixDouble)
 *   gdp_matrix_scale      m2 == m1: m1 *= d           (LinAlg_ProdMatrixDouble)
 * Patterns may differ: the result's pattern is their union, an entry of m2
 * alone holds (d *) its value.  gdp_matrix_size: LinAlg_GetMatrixSize. */
int gdp_matrix_add(gdp_matrix *m1, gdp_matrix *m2, gdp_matrix *m3);
int gdp_matrix_add_scaled(gdp_matrix *m1, gdp_matrix *m2, double d, gdp_matrix *m3);
int gdp_matrix_scale(gdp_matrix *m1, double d, gdp_matrix *m2);
int gdp_matrix_size(const gdp_matrix *A, int *n, int *m);
int gdp_matrix_create(int n, int m, gdp_matrix **A);          /* LinAlg_CreateMatrix  */
int gdp_matrix_destroy(gdp_matrix *A);                        /* LinAlg_DestroyMatrix */
/* LinAlg_ZeroMatrix: keeps a device pattern (values zeroed) when one exists;
 * staged host entries are discarded. */
int gdp_matrix_zero(gdp_matrix *A);
/* LinAlg_AddDoubleInMatrix (LinAlg_SPARSKIT.cpp:392-400): row i, column j.
 * Entries are staged on the host; duplicates are summed in call order by
 * gdp_matrix_assemble (add_matrix_double, Sparskit.cpp:809-862). */
int gdp_matrix_add_double(gdp_matrix *A, int i, int j, double d);
int gdp_matrix_add_coo(gdp_matrix *A, long n, const int *i, const int *j,
                       const double *d);
/* LinAlg_AssembleMatrix: builds the sorted CSR (csr_format, Sparskit.cpp:2186)
 * from the staged entries and uploads it. */
int gdp_matrix_assemble(gdp_matrix *A);
int gdp_matrix_set_csr(gdp_matrix *A, int n, const int *ia, const int *ja,
                       const double *a);
int gdp_matrix_nnz(const gdp_matrix *A, long *nnz);
int gdp_matrix_get_csr(const gdp_matrix *A, int *ia, int *ja, double *a);
/* LinAlg_ProdMatrixVector: y = A x on the device (CSR SpMV, FP64). */
int gdp_prod_matrix_vector(gdp_matrix *A, const gdp_vector *x, gdp_vector *y);

/* Measurement helper, no counterpart in the reference (for benchmarks and
 * profiling): `reps` back-to-back launches of y = A x on the library stream,
 * bracketed by HIP events; *avg_ms = average kernel time per launch. */
int gdp_bench_spmv(gdp_matrix *A, const gdp_vector *x, gdp_vector *y, int reps,
                   double *avg_ms);
/* Measurement helpers, no counterpart in the reference.  The matrix's SpMV
 * layout: padded entry count and whether its column indices are stored as
 * 16-bit offsets (1) or 32-bit (0). */
int gdp_matrix_sell_info(gdp_matrix *A, long *nnz_pad, int *compressed);
/* `reps` launches of the SpMV on the solver's renumbered (and scaled) copy of
 * the last matrix it solved -- the operator of its Krylov iteration --
 * bracketed by HIP events; *avg_ms per launch, *compressed as above. */
int gdp_solver_bench_spmv(gdp_solver *s, int reps, double *avg_ms, int *compressed);
/* The lusol step of solve_matrix (ilut.F:2297-2347, called from the
 * reverse-communication loop of Sparskit.cpp:556-567): out = (LU)^{-1} in with
 * the ILU(0) factors of the solver's last solve, in its internal numbering
 * (RCMK-renumbered, scaled system), `reps` times in a row; *avg_ms (may be
 * null) = average time per application.  For tests and measurements. */
int gdp_solver_prec_apply(gdp_solver *s, const gdp_vector *in, gdp_vector *out, int reps,
                          double *avg_ms);

/* ---- solve ----------------------------------------------------------------- */
/* LinAlg_Solve (LinAlg_SPARSKIT.cpp:572 -> solve_matrix, Sparskit.cpp:108):
 * scaling, RCMK renumbering, ILU(0) and CG (Algorithm 1, iters.f:379-562) /
 * BCGSTAB (5, iters.f:1032-1306) / GMRES (8) / FGMRES (9) on the device,
 * x0 = 0, stop on the estimated residual <= Stopping_Test * |r0|. */
int gdp_solve(gdp_matrix *A, gdp_vector *b, gdp_solver *s, gdp_vector *x);
/* LinAlg_SolveAgain (:578): reuses permutation and factors of the last solve
 * with the same matrix. */
int gdp_solve_again(gdp_matrix *A, gdp_vector *b, gdp_solver *s, gdp_vector *x);

/* ---- batched Galerkin assembly ---------------------------------------------- */
/* Element types: Gmsh codes 1 line, 2 triangle, 4 tetrahedron. */
#define GDP_LINE 1
#define GDP_TRI 2
#define GDP_TET 4
/* Basis functions: BF_Node (BF_Node.cpp:28) on NodesOf, BF_Edge (BF_Edge.cpp:23)
 * on EdgesOf, BF_PerpendicularEdge (BF_Perpendicular.cpp:33-41: (0, 0, BF_Node),
 * curl BF_CurlPerpendicularEdge :70-78) on NodesOf of triangles in the plane.
 * On a perpendicular-edge space GDP_T_CURL_CURL, GDP_T_EDGE_MASS and
 * GDP_T_SRC_EDGE take the 2D forms Form2P / Form1P (Get_Geometry.cpp:1929-1947). */
#define GDP_BF_NODE 0
#define GDP_BF_EDGE 1
#define GDP_BF_PERP_EDGE 2
/* Canonical terms recognised by Cal_InitGalerkinTermOfFemEquation
 * (Cal_GalerkinTermOfFemEquation.cpp:39-316) and supported on the device: */
#define GDP_T_GRAD_GRAD 1  /* [ c * Dof{d u}, {d v} ]: c scalar or 3x3 tensor  */
#define GDP_T_CURL_CURL 2  /* [ c * Dof{d a}, {d a} ]: c scalar or 3x3 tensor  */
#define GDP_T_EDGE_MASS 3  /* [ c * Dof{a}, {a} ], Whitney edges (tets)        */
#define GDP_T_NODE_MASS 4  /* [ c * Dof{u}, {u} ]                              */
#define GDP_T_SRC_NODE 10  /* [ f, {u} ]: no Dof{}, f scalar                    */
#define GDP_T_SRC_EDGE 11  /* [ f, {a} ]: no Dof{}, f vector                    */

typedef struct {
  int nnodes;
  const double *xyz;  /* 3 per node                                          */
  int nelem;          /* elements in ascending element number               */
  const int *etype;   /* nelem                                               */
  const int *region;  /* nelem, physical region ids (>= 0)                   */
  const int *conn;    /* 4 per element, 1-based node numbers                 */
  const int *edges;   /* 6 per element, signed edge numbers (NULL: no edges) */
  int nedges;         /* number of edges (0 if edges == NULL)                */
} gdp_mesh_desc;

typedef struct {
  int bf;                    /* GDP_BF_NODE / GDP_BF_EDGE / GDP_BF_PERP_EDGE  */
  int numtype;               /* BasisFunction.Num (DOF key, DofData.cpp:36)   */
  int nsupport;
  const int *support;        /* support regions                              */
  int nfixed;                /* Dirichlet-type constraint                     */
  const int *fixed_entity;   /* node numbers (1-based) or edge numbers        */
  const double *fixed_value;
} gdp_space_desc;

typedef struct {
  int kind;                  /* GDP_T_*                                       */
  int equ, dof;              /* space indices; dof = -1 for source terms      */
  int nregions;
  const int *regions;        /* In ...                                        */
  int npoints;               /* Gauss points (tri 1/3/4, tet 1/4)             */
  int jac_sur;               /* 1: Jacobian Sur (triangles in 3D)             */
  int ncoef;                 /* 1 scalar, 3 vector, 9 tensor                  */
  int coef_nreg;             /* table rows = region ids 0..coef_nreg-1        */
  const double *coef;        /* coef[r*ncoef + c], sign included              */
} gdp_term_desc;

typedef struct gdp_fe gdp_fe;
/* PRE pass for one formulation (SolvingAnalyse.cpp:445 -> Treatment_Formulation
 * in STATUS_PRE): DOF numbering (DofData.cpp:1279-1323), CSR pattern
 * (every DOF pair sharing an element, stored zeros included), gather plan and
 * device upload of mesh and coefficient tables.  Fails with
 * GDP_E_UNSUPPORTED for terms the device path does not cover. */
int gdp_fe_create(const gdp_mesh_desc *mesh, int nspaces,
                  const gdp_space_desc *spaces, int nterms,
                  const gdp_term_desc *terms, gdp_fe **fe);
int gdp_fe_destroy(gdp_fe *fe);
int gdp_fe_num_dofs(const gdp_fe *fe, int *N, long *nnz);
/* dofnum[entity] for entities 0..nmax (index 0 unused): > 0 unknown number
 * (1-based), < 0 -(index into fixed_entity + 1), 0 not a DOF. */
int gdp_fe_dof_table(const gdp_fe *fe, int space, int *dofnum, int nmax);
/* The formulation's CSR pattern (rows = equation DOFs, sorted columns):
 * ia (N+1) and ja (nnz); host data, available without a device. */
int gdp_fe_pattern(const gdp_fe *fe, int *ia, int *ja);
/* Dof_WriteFilePRE (DofData.cpp:363-460): the $DofData block of a .pre file
 * for the formulation's DOFs (DofData number num, its resolution and system
 * indices and function space indices as GetDP holds them), one line per DOF
 * in (NumType, Entity) order, in Dof_WriteDofPRE's formats.  Host data,
 * available without a device. */
int gdp_fe_write_pre(const gdp_fe *fe, FILE *f, int num, int resolution_index, int system_index, int nfs,
                     const int *fs_index);
/* Generate[A] (SolvingOperations.cpp:276-500 -> per element
 * Cal_GalerkinTermOfFemEquation + Dof_AssembleInMat): zeroes A and b, then
 * assembles every term of every element on the device, summing each entry in
 * ascending element order as the reference does.  A takes the formulation's
 * pattern on first use. */
int gdp_fe_generate(gdp_fe *fe, gdp_matrix *A, gdp_vector *b);
/* ---- post-processing (PostOperation Print[...] of local and integral
 * quantities, Cal_PostQuantity.cpp:29-247 -> Pos_FemInterpolation,
 * Pos_FemInterpolation.cpp:226-330) -------------------------------------- */
#define GDP_PQ_VALUE 1        /* {u} at each vertex: nodal scalar, Whitney     */
                              /* vector, or (0,0,phi) on perpendicular edges  */
#define GDP_PQ_D 2            /* {d u}: gradient / curl / 2D curl, per element */
#define GDP_PQ_COEF_D 3       /* coef[] * {d u} (scalar or 3x3 tensor)         */
#define GDP_PQ_INT_COEF_D2 4  /* Integral { [ coef[] * SquNorm[{d u}] ] }      */
typedef struct {
  int kind;                   /* GDP_PQ_*                                      */
  int space;                  /* function space index                          */
  int nregions;
  const int *regions;         /* OnElementsOf / In                             */
  int ncoef;                  /* COEF_D: 1 or 9; INT_COEF_D2: 1                */
  int coef_nreg;              /* table rows = region ids 0..coef_nreg-1        */
  const double *coef;         /* coef[r*ncoef + c]                             */
  int npoints;                /* INT_COEF_D2: Gauss points (tri 1/3/4, tet 1/4)*/
} gdp_post_desc;
/* Evaluates quantity q of the solution x (original DOF numbering, fixed DOFs
 * taking their constraint values) on the elements of q->regions, in element
 * order; *nsel = their count.  out (may be NULL): for VALUE, D and COEF_D,
 * 12 values per element (4 vertices x 3 components, zero-padded; vertices in
 * connectivity order); for INT_COEF_D2 the element integral (Gauss rule as in
 * the integral quantity's Integration) and *total their sum (Print[q[Region],
 * OnGlobal]).  *kernel_ms (may be NULL): device time of the evaluation. */
int gdp_fe_post(gdp_fe *fe, const gdp_vector *x, const gdp_post_desc *q, int *nsel,
                double *out, long out_len, double *total, double *kernel_ms);
/* Device time of the last gdp_fe_generate (ms) and of its kernels. */
int gdp_fe_last_timing(const gdp_fe *fe, double *ms_total, double *ms_geometry,
                       double *ms_gather);

/* ---- mesh files ------------------------------------------------------------ */
/* Gmsh MSH 2.x, ASCII or binary (Geo_ReadFile, GeoData.cpp:573-1028): nodes
 * and elements as the reference keeps them -- elements sorted by number,
 * region = first tag (1 without tags in ASCII files, 0 in binary ones).
 * Node numbers are mapped to 1..nnodes in ascending order (node_num gives the
 * originals); conn holds those indices, 4 per element (zero-padded).  Point
 * elements are counted (npoints) and not kept; types other than points,
 * lines, triangles and tetrahedra are rejected.  Host code, no device. */
typedef struct gdp_mesh_data gdp_mesh_data;
int gdp_mesh_read(const char *path, gdp_mesh_data **mesh);
int gdp_mesh_info(const gdp_mesh_data *mesh, int *nnodes, int *nelem, int *npoints, int *binary);
int gdp_mesh_get(const gdp_mesh_data *mesh, double *xyz, int *node_num, int *elem_num, int *etype,
                 int *region, int *elementary, int *conn);
int gdp_mesh_free(gdp_mesh_data *mesh);
/* Writes MSH 2.2 (ASCII with %.17g coordinates, or binary): nodes 1..nnodes,
 * elements numbered 1..nelem with tags (region, region). */
int gdp_mesh_write(const char *path, int binary, int nnodes, const double *xyz, int nelem,
                   const int *etype, const int *region, const int *conn);

/* ---- pre-processing of mesh entities and patterns -------------------------- */
/* Edge creation (GeoEntity.cpp:437-490): visiting elements in `visit` order,
 * each new edge (sorted node pair) gets the next number; signed by local
 * orientation (+ when it runs from the lower to the higher node number).
 * edges_out: 6 per element, edge_nodes: 2 per edge.  Returns #edges or < 0.
 * Computed on the device (sort-based, same numbers) once the library is
 * initialised unless GDP_HOST_PREP is set; on the host otherwise.  The other
 * functions of this section run on the host. */
int gdp_number_edges(int nelem, const int *etype, const int *conn, int nvisit,
                     const int *visit, int *edges_out, int *edge_nodes,
                     int max_edges);
/* RCMK renumbering (cmkreord.f:3-62 with perphn/BFS of reordering.f):
 * perm[old] = new, rperm[new] = old (0-based).  Returns #levels or < 0. */
int gdp_rcmk(int n, const int *ia, const int *ja, int *perm, int *rperm);
/* ILUT (pivot 0, ilut.F:52-396) or ILUTP (pivot 1, ilut.F:397-806 with
 * mbloc = n) factors of the 0-based CSR matrix (ia, ja, a), on the host, as
 * gdp_solve computes them for Preconditioner 1 / 2; factor storage FP32
 * unless f64 (the values are returned in double, exactly).  Row i occupies
 * [ptr[i], ptr[i+1]): its L entries, then at dpos[i] the inverted pivot, then
 * its U entries, each part in the order the triangular sweeps accumulate it;
 * columns in the permuted numbering, newpos[c] = permuted number of column c
 * (identity for ILUT).  *nnz_out = number of entries; when it exceeds
 * capacity only ptr is filled and GDP_E_NOMEM is returned (the reference's
 * iwk overflow, ierr -2/-3): call again with capacity >= *nnz_out. */
int gdp_ilut(int n, const int *ia, const int *ja, const double *a, int lfil, double droptol,
             double permtol, int pivot, int f64, long capacity, int *ptr, int *dpos, int *col,
             double *val, int *newpos, long *nnz_out);

/* ---- batched solves (no single counterpart in the reference) ----------------
 * Many independent systems solved as the reference solves one (Algorithm 8,
 * Preconditioner 6: solve_matrix, Sparskit.cpp:108-634, per instance).  The
 * systems are the diagonal blocks of one matrix -- instance i owns rows
 * [row_offsets[i], row_offsets[i+1]) -- typically assembled as one
 * formulation over concatenated meshes.  One persistent kernel per launch
 * configuration runs whole instances per workgroup, largest first; the
 * assembly (when fe is given), renumbering of values, ILU(0) factorisation,
 * the solves and the unpermutation are one captured hipGraph, recorded on
 * the first gdp_batch_run and replayed after.
 * config (may be NULL): 3 ints per instance -- workgroup size (256, 512,
 * 1024), orthogonalisation (0 modified, 1 classical twice), ILU ordering
 * (0 as Renumbering_Technique, 1 multicolour on top).  Krylov_Size <= 32. */
typedef struct gdp_batch gdp_batch;
int gdp_batch_create(const gdp_solver *params, int ninst, const int *row_offsets,
                     const int *config, gdp_batch **out);
int gdp_batch_destroy(gdp_batch *batch);
int gdp_batch_run(gdp_batch *batch, gdp_fe *fe, gdp_matrix *A, gdp_vector *b, gdp_vector *x);
/* Per instance: iterations, status (0 converged, -1 iteration limit), |r0|,
 * estimated final residual norm, start/end on the device clock in us from
 * the earliest start; graph_ms: device time of the last run. */
int gdp_batch_info(const gdp_batch *batch, int *its, int *status, double *res0, double *res,
                   double *t_begin_us, double *t_end_us, double *graph_ms);

/* ---- learned preconditioner (no counterpart in the reference) --------------
 * A 4-level convolutional U-Net (16-32-64-128 channels, 3x3 kernels, ReLU,
 * 2x2 average pooling, nearest upsampling + skip concatenation, 1x1 output)
 * on an H x W grid of nodal values, BF16 operands with FP32 accumulation on
 * the MFMA units, FP32 master weights trained with Adam.  Input channels:
 * the residual (normalised to unit RMS) and a per-pixel coefficient feature
 * (log nu); output: the correction, in the same normalisation.  Bound to a
 * solver with gdp_unet_attach it is the right preconditioner of flexible
 * GMRES (Algorithm 9, iters.f:2522-2807), FP64 or FP32 (Precision 1). */
typedef struct gdp_unet gdp_unet;
/* H, W multiples of 8; batch = largest batch of train/forward calls. */
int gdp_unet_create(int H, int W, int batch, unsigned seed, gdp_unet **out);
/* The same network on a 3D grid of D x H x W voxels (each side a multiple of
 * 8): 3x3x3 convolutions, 2x2x2 pooling and upsampling; parameters in the
 * layout [cout][27][cin], taps ordered (dz, dy, dx) row-major.  Every other
 * gdp_unet_* call works on either kind, with voxels in (z, y, x) order. */
int gdp_unet_create_3d(int D, int H, int W, int batch, unsigned seed, gdp_unet **out);
int gdp_unet_destroy(gdp_unet *net);
int gdp_unet_num_params(const gdp_unet *net, long *count);
/* Parameters as one FP32 array: per convolution weights [cout][3][3][cin]
 * then bias [cout]; then the output weights [16] and bias [1]. */
int gdp_unet_get_params(gdp_unet *net, float *host);
int gdp_unet_set_params(gdp_unet *net, const float *host);
/* Gradients of the last gdp_unet_train_step, in the parameter layout. */
int gdp_unet_get_grads(gdp_unet *net, float *host);
/* r, feat: batch x H x W (residuals already normalised; mask: H x W). */
int gdp_unet_forward(gdp_unet *net, int batch, const float *r, const float *feat, float *z);
/* One Adam step (lr = 0: gradients only) on the loss
 * |z m / |z m| - t|^2 per sample (t unit-norm target on the mask m). */
int gdp_unet_train_step(gdp_unet *net, int batch, const float *r, const float *feat,
                        const float *target, const float *mask, float lr, float *loss);
/* Tensor t of the last forward pass as FP32 NHWC (t = 0 input, 1..14 the
 * convolution outputs, 15 the network output); host = NULL returns the count. */
int gdp_unet_get_activation(gdp_unet *net, int t, int batch, float *host, long *count);
/* pix2dof[H*W]: the DOF (0-based, original numbering) at each pixel or -1;
 * every DOF must appear.  feat[H*W]: coefficient feature per pixel.
 * dscale[n_dof] (may be NULL): the network works on s*r and returns e/s per
 * DOF, s = 1/sqrt(diag A) making the operator unit-diagonal. */
int gdp_unet_attach(gdp_unet *net, gdp_solver *solver, int n_dof, const int *pix2dof,
                    const float *feat, const double *dscale);
int gdp_unet_detach(gdp_solver *solver);
/* Preconditioner options, bit 0: apply the odd part (N(r) - N(-r))/2 of the
 * network (batch >= 2 at creation).  With Preconditioner 6 on the solver the
 * preconditioner is two-stage: z = a N(v) + (LU)^{-1}(v - a A N(v)), a = 1,
 * or with bit 1 set the least-squares scale a = <v, A N(v)> / |A N(v)|^2. */
int gdp_unet_set_mode(gdp_unet *net, int flags);

#ifdef __cplusplus
}
#endif
#endif /* GDP_HIP_H */
