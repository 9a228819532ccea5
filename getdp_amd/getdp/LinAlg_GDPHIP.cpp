// GetDP's LinAlg interface (src/kernel/LinAlg.h:90-213) on the MI355X
// library: every function of the Sparskit backend (LinAlg_SPARSKIT.cpp), with
// the same argument rules and messages, forwarding to the C-ABI of gdp_hip.h.
// Compiled into GetDP in place of LinAlg_SPARSKIT.cpp, with HAVE_GDP_HIP and
// the LinAlg_GDPHIP.h branch of LinAlg.h.
//
// Vectors and matrices live on the device.  Entry-wise vector accessors go
// through a host mirror (LinAlg_GDPHIP.h), as GetDP's post-processing reads
// solution entries one by one.  Functions that the Sparskit backend leaves
// unimplemented ("not yet implemented") keep its error.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "LinAlg.h"
#include "Message.h"

extern char *Name_Path;

#if defined(HAVE_GDP_HIP)

// command-line solver options ("-Name value"), the solver file given with
// -solver, the default file name (LinAlg_SPARSKIT.cpp:23-66)
static const char *Name_SolverFile = NULL, *Name_DefaultSolverFile = "solver.par";
static std::vector<std::string> SolverOptions;

// ---- status mapping --------------------------------------------------------------

static bool ok(int rc, const char *what)
{
  if(rc == GDP_OK) return true;
  Message::Error("%s: %s", what, gdp_last_error());
  return false;
}

// ---- host mirror of a vector --------------------------------------------------------

static void mirror_alloc(gVector *V)
{
  if(!V->mirror) V->mirror = (double *)calloc(V->N > 0 ? V->N : 1, sizeof(double));
}

// entries as the device holds them (after pending additions)
static void to_host(gVector *V)
{
  mirror_alloc(V);
  if(V->mirror_state == 0 && V->N) ok(gdp_vector_get(V->V, V->mirror, V->N), "LinAlg (vector read)");
  if(V->mirror_state == 0) V->mirror_state = 1;
}

// entry-wise writes made on the mirror go to the device first
static void to_device(gVector *V)
{
  if(V->mirror_state == 2) {
    ok(gdp_vector_set(V->V, V->mirror, V->N), "LinAlg (vector write)");
    V->mirror_state = 1;
  }
}

// before a device operation that changes V
static void device_write(gVector *V)
{
  to_device(V);
  V->mirror_state = 0;
}

static void set_entry(gVector *V, int i, double d)
{
  to_host(V);
  V->mirror[i] = d;
  V->mirror_state = 2;
}

// ---- solver ---------------------------------------------------------------------------

void LinAlg_InitializeSolver(int *sargc, char ***sargv)
{
  int i = 1, argc = *sargc;
  char **argv = *sargv;
  SolverOptions.clear();
  while(i < argc) {
    if(argv[i][0] == '-') {
      if(!strcmp(argv[i] + 1, "solver") || !strcmp(argv[i] + 1, "s")) {
        i++;
        if(i < argc && argv[i][0] != '-') Name_SolverFile = argv[i++];
        else Message::Error("Missing file name");
      }
      else if(!strcmp(argv[i] + 1, "slepc")) {
        i++;
      }
      else {
        i++;
        if(i < argc && argv[i][0] != '-') {
          SolverOptions.push_back(argv[i - 1] + 1);
          SolverOptions.push_back(argv[i]);
          i++;
        }
        else
          Message::Error("Missing number");
      }
    }
    else
      Message::Info("Unknown option: '%s'", argv[i++]);
  }
  ok(gdp_init(0), "LinAlg_InitializeSolver");
}

void LinAlg_FinalizeSolver() { gdp_finalize(); }

void LinAlg_SetCommSelf() {}

void LinAlg_SetCommWorld() {}

void LinAlg_CreateSolver(gSolver *Solver, const char *SolverDataFileName)
{
  std::string file = Name_Path ? Name_Path : "";
  if(SolverDataFileName) {
    if(SolverDataFileName[0] == '/' || SolverDataFileName[0] == '\\') file = SolverDataFileName;
    else file += SolverDataFileName;
  }
  else if(Name_SolverFile)
    file = Name_SolverFile;
  else
    file += Name_DefaultSolverFile;
  Message::Info("Loading parameter file '%s'", file.c_str());
  Solver->S = NULL;
  if(!ok(gdp_solver_create(file.c_str(), &Solver->S), "LinAlg_CreateSolver")) return;
  for(size_t k = 0; k + 1 < SolverOptions.size(); k += 2)
    ok(gdp_solver_set(Solver->S, SolverOptions[k].c_str(), SolverOptions[k + 1].c_str()), "solver option");
}

void LinAlg_SetGlobalSolverOptions(const std::string &opt) {}

void LinAlg_DestroySolver(gSolver *Solver)
{
  gdp_solver_destroy(Solver->S);
  Solver->S = NULL;
}

// ---- creation, copies ------------------------------------------------------------------

void LinAlg_CreateVector(gVector *V, gSolver *Solver, int n)
{
  V->N = n;
  V->V = NULL;
  V->mirror = NULL;
  V->mirror_state = 1;  // a new vector is zero on both sides
  ok(gdp_vector_create(n, &V->V), "LinAlg_CreateVector");
  mirror_alloc(V);
}

void LinAlg_CreateMatrix(gMatrix *M, gSolver *Solver, int n, int m)
{
  M->M = NULL;
  ok(gdp_matrix_create(n, m, &M->M), "LinAlg_CreateMatrix");
}

void LinAlg_DestroyVector(gVector *V)
{
  gdp_vector_destroy(V->V);
  free(V->mirror);
  V->V = NULL;
  V->mirror = NULL;
}

void LinAlg_DestroyMatrix(gMatrix *M)
{
  gdp_matrix_destroy(M->M);
  M->M = NULL;
}

void LinAlg_CopyScalar(gScalar *S1, gScalar *S2) { S2->s = S1->s; }

void LinAlg_CopyVector(gVector *V1, gVector *V2)
{
  to_device(V1);
  device_write(V2);
  ok(gdp_vector_copy(V1->V, V2->V), "LinAlg_CopyVector");
}

void LinAlg_CopyMatrix(gMatrix *M1, gMatrix *M2) { Message::Error("'LinAlg_CopyMatrix' not yet implemented"); }

void LinAlg_SwapVector(gVector *V1, gVector *V2)
{
  if(V1->N != V2->N) {
    Message::Error("Cannot swap vectors of different size");
    return;
  }
  device_write(V1);
  device_write(V2);
  ok(gdp_vector_swap(V1->V, V2->V), "LinAlg_SwapVector");
}

// ---- zero --------------------------------------------------------------------------------

void LinAlg_ZeroScalar(gScalar *S) { S->s = 0.; }

void LinAlg_ZeroVector(gVector *V)
{
  V->mirror_state = 0;
  ok(gdp_vector_zero(V->V), "LinAlg_ZeroVector");
}

// The Sparskit backend also adds a 0 on every diagonal entry (its product
// fails on empty rows); the device pattern keeps every row, and Generate[A]
// gives the formulation's pattern on first use.
void LinAlg_ZeroMatrix(gMatrix *M) { ok(gdp_matrix_zero(M->M), "LinAlg_ZeroMatrix"); }

// ---- text and binary I/O ------------------------------------------------------------------

void LinAlg_ScanScalar(FILE *file, gScalar *S)
{
  if(fscanf(file, "%lf", &S->s) != 1) Message::Error("Could not read a scalar");
}

void LinAlg_ScanVector(FILE *file, gVector *V)
{
  V->mirror_state = 0;
  ok(gdp_vector_scan(V->V, file), "LinAlg_ScanVector");
}

void LinAlg_ScanMatrix(FILE *file, gMatrix *M) { Message::Error("'LinAlg_ScanMatrix' not yet implemented"); }

void LinAlg_ReadScalar(FILE *file, gScalar *S) { Message::Error("'LinAlg_ReadScalar' not yet implemented"); }

void LinAlg_ReadVector(FILE *file, gVector *V)
{
  V->mirror_state = 0;
  ok(gdp_vector_read(V->V, file), "LinAlg_ReadVector");
}

void LinAlg_ReadMatrix(FILE *file, gMatrix *M) { Message::Error("'LinAlg_ReadMatrix' not yet implemented"); }

void LinAlg_PrintScalar(FILE *file, gScalar *S) { fprintf(file, "%.16g", S->s); }

void LinAlg_PrintVector(FILE *file, gVector *V, bool matlab, const char *fileName, const char *varName)
{
  if(matlab) Message::Error("Matlab output not available for this vector");
  to_device(V);
  ok(gdp_vector_print(V->V, file), "LinAlg_PrintVector");
}

// "row  col  value" per entry in CSR order, 1-based (the KUL style of
// formatted_write_matrix, Sparskit.cpp:1329-1340)
static bool matrix_host(gMatrix *M, std::vector<int> &ia, std::vector<int> &ja, std::vector<double> &a)
{
  int n = 0, m = 0;
  long nnz = 0;
  if(!ok(gdp_matrix_size(M->M, &n, &m), "matrix size") || !ok(gdp_matrix_nnz(M->M, &nnz), "matrix size"))
    return false;
  ia.resize((size_t)n + 1);
  ja.resize((size_t)nnz);
  a.resize((size_t)nnz);
  return ok(gdp_matrix_get_csr(M->M, ia.data(), ja.data(), a.data()), "matrix read");
}

void LinAlg_PrintMatrix(FILE *file, gMatrix *M, bool matlab, const char *fileName, const char *varName)
{
  if(matlab) Message::Error("Matlab output not available for this matrix");
  std::vector<int> ia, ja;
  std::vector<double> a;
  if(!matrix_host(M, ia, ja, a)) return;
  if(ia.size() <= 1) {
    Message::Warning("No element in matrix");
    return;
  }
  for(size_t i = 0; i + 1 < ia.size(); i++)
    for(int k = ia[i]; k < ia[i + 1]; k++) fprintf(file, "%d  %d  %.16g\n", (int)i + 1, ja[k] + 1, a[k]);
}

void LinAlg_WriteScalar(FILE *file, gScalar *S) { Message::Error("'LinAlg_WriteScalar' not yet implemented"); }

void LinAlg_WriteVector(FILE *file, gVector *V)
{
  to_device(V);
  ok(gdp_vector_write(V->V, file), "LinAlg_WriteVector");
}

// The Sparskit backend dumps its own storage to A.mat (binary_write_matrix,
// Sparskit.cpp:1225-1270).  Here: "n nnz" on a text line, then the 0-based
// CSR arrays ia (n+1 int), ja (nnz int) and a (nnz double), binary.
void LinAlg_WriteMatrix(FILE *file, gMatrix *M)
{
  std::vector<int> ia, ja;
  std::vector<double> a;
  if(!matrix_host(M, ia, ja, a)) return;
  FILE *f = fopen("A.mat", "wb");
  if(!f) {
    Message::Error("Could not open file 'A.mat'");
    return;
  }
  fprintf(f, "%d %ld\n", (int)ia.size() - 1, (long)ja.size());
  fwrite(ia.data(), sizeof(int), ia.size(), f);
  fwrite(ja.data(), sizeof(int), ja.size(), f);
  fwrite(a.data(), sizeof(double), a.size(), f);
  fclose(f);
}

// ---- sizes and entries ----------------------------------------------------------------------

void LinAlg_GetVectorSize(gVector *V, int *i) { *i = V->N; }

void LinAlg_GetLocalVectorRange(gVector *V, int *low, int *high)
{
  *low = 0;
  *high = V->N;
}

void LinAlg_GetMatrixSize(gMatrix *M, int *i, int *j)
{
  int m = 0;
  ok(gdp_matrix_size(M->M, i, &m), "LinAlg_GetMatrixSize");
  *j = *i;
}

void LinAlg_GetLocalMatrixRange(gMatrix *M, int *low, int *high)
{
  int m = 0;
  *low = 0;
  ok(gdp_matrix_size(M->M, high, &m), "LinAlg_GetLocalMatrixRange");
}

void LinAlg_GetDoubleInScalar(double *d, gScalar *S) { *d = S->s; }

void LinAlg_GetComplexInScalar(double *d1, double *d2, gScalar *S)
{
  Message::Error("'LinAlg_GetComplexInScalar' not available with this Solver");
}

void LinAlg_GetScalarInVector(gScalar *S, gVector *V, int i)
{
  to_host(V);
  S->s = V->mirror[i];
}

void LinAlg_GetDoubleInVector(double *d, gVector *V, int i)
{
  to_host(V);
  *d = V->mirror[i];
}

void LinAlg_GetAbsDoubleInVector(double *d, gVector *V, int i)
{
  to_host(V);
  *d = fabs(V->mirror[i]);
}

void LinAlg_GetComplexInVector(double *d1, double *d2, gVector *V, int i, int j)
{
  to_host(V);
  *d1 = V->mirror[i];
  *d2 = V->mirror[j];
}

void LinAlg_GetScalarInMatrix(gScalar *S, gMatrix *M, int i, int j)
{
  Message::Error("'LinAlg_GetScalarInMatrix' not yet implemented");
}

// get_element_in_matrix (Sparskit.cpp): the entry, 0 outside the pattern
void LinAlg_GetDoubleInMatrix(double *d, gMatrix *M, int i, int j)
{
  std::vector<int> ia, ja;
  std::vector<double> a;
  *d = 0.;
  if(!matrix_host(M, ia, ja, a)) return;
  for(int k = ia[i]; k < ia[i + 1]; k++)
    if(ja[k] == j) *d = a[k];
}

void LinAlg_GetComplexInMatrix(double *d1, double *d2, gMatrix *M, int i, int j, int k, int l)
{
  Message::Error("'LinAlg_GetComplexInMatrix' not yet implemented");
}

// get_column_in_matrix (Sparskit.cpp): column col of M into V1
void LinAlg_GetColumnInMatrix(gMatrix *M, int col, gVector *V1)
{
  std::vector<int> ia, ja;
  std::vector<double> a;
  if(!matrix_host(M, ia, ja, a)) return;
  std::vector<double> c(V1->N, 0.);
  for(size_t i = 0; i + 1 < ia.size() && (int)i < V1->N; i++)
    for(int k = ia[i]; k < ia[i + 1]; k++)
      if(ja[k] == col) c[i] = a[k];
  V1->mirror_state = 0;
  ok(gdp_vector_set(V1->V, c.data(), V1->N), "LinAlg_GetColumnInMatrix");
}

void LinAlg_SetScalar(gScalar *S, double *d) { S->s = d[0]; }

void LinAlg_SetVector(gVector *V, double *v)
{
  V->mirror_state = 0;
  ok(gdp_vector_set_all(V->V, *v), "LinAlg_SetVector");
}

void LinAlg_SetScalarInVector(gScalar *S, gVector *V, int i) { set_entry(V, i, S->s); }

void LinAlg_SetDoubleInVector(double d, gVector *V, int i) { set_entry(V, i, d); }

void LinAlg_SetComplexInVector(double d1, double d2, gVector *V, int i, int j)
{
  set_entry(V, i, d1);
  set_entry(V, j, d2);
}

void LinAlg_SetScalarInMatrix(gScalar *S, gMatrix *M, int i, int j)
{
  Message::Error("'LinAlg_SetScalarInMatrix' not yet implemented");
}

void LinAlg_SetDoubleInMatrix(double d, gMatrix *M, int i, int j)
{
  Message::Error("'LinAlg_SetDoubleInMatrix' not yet implemented");
}

void LinAlg_SetComplexInMatrix(double d1, double d2, gMatrix *M, int i, int j, int k, int l)
{
  Message::Error("'LinAlg_SetComplexInMatrix' not yet implemented");
}

// ---- scalar arithmetic -------------------------------------------------------------------------

void LinAlg_AddScalarScalar(gScalar *S1, gScalar *S2, gScalar *S3) { S3->s = S1->s + S2->s; }

void LinAlg_SubScalarScalar(gScalar *S1, gScalar *S2, gScalar *S3) { S3->s = S1->s - S2->s; }

void LinAlg_ProdScalarScalar(gScalar *S1, gScalar *S2, gScalar *S3) { S3->s = S1->s * S2->s; }

void LinAlg_ProdScalarDouble(gScalar *S1, double d, gScalar *S2) { S2->s = S1->s * d; }

void LinAlg_ProdScalarComplex(gScalar *S, double d1, double d2, double *d3, double *d4)
{
  *d3 = S->s * d1;
  *d4 = S->s * d2;
}

void LinAlg_DivScalarScalar(gScalar *S1, gScalar *S2, gScalar *S3) { S3->s = S1->s / S2->s; }

void LinAlg_DivScalarDouble(gScalar *S1, double d, gScalar *S2) { S2->s = S1->s / d; }

// ---- assembly into vectors and matrices ------------------------------------------------------------

// DummyDof filter of LinAlg_SPARSKIT.cpp:335-400: entries of dummy DOFs are
// dropped, diagonal matrix entries excepted
static bool dummy(int i)
{
  int *dd = GDP_CurrentDummyDof();
  return dd && dd[i] == 1;
}

void LinAlg_DummyVector(gVector *V)
{
  int *dd = GDP_CurrentDummyDof();
  if(!dd) return;
  device_write(V);
  ok(gdp_vector_dummy(V->V, dd), "LinAlg_DummyVector");
}

static void add_entry(gVector *V, int i, double d)
{
  if(V->mirror_state == 2) {  // the mirror holds entry-wise writes: add there, in order
    V->mirror[i] += d;
    return;
  }
  V->mirror_state = 0;
  ok(gdp_vector_add_double(V->V, i, d), "LinAlg_AddDoubleInVector");
}

void LinAlg_AddScalarInVector(gScalar *S, gVector *V, int i)
{
  if(dummy(i)) return;
  add_entry(V, i, S->s);
}

void LinAlg_AddDoubleInVector(double d, gVector *V, int i)
{
  if(dummy(i)) return;
  add_entry(V, i, d);
}

void LinAlg_AddComplexInVector(double d1, double d2, gVector *V, int i, int j)
{
  if(!dummy(i)) add_entry(V, i, d1);
  if(!dummy(j)) add_entry(V, j, d2);
}

void LinAlg_AddScalarInMatrix(gScalar *S, gMatrix *M, int i, int j)
{
  if((dummy(i) || dummy(j)) && i != j) return;
  ok(gdp_matrix_add_double(M->M, i, j, S->s), "LinAlg_AddScalarInMatrix");
}

void LinAlg_AddDoubleInMatrix(double d, gMatrix *M, int i, int j)
{
  if((dummy(i) || dummy(j)) && i != j) return;
  ok(gdp_matrix_add_double(M->M, i, j, d), "LinAlg_AddDoubleInMatrix");
}

// the real 2x2 block form of a complex entry (LinAlg_SPARSKIT.cpp:402-413)
void LinAlg_AddComplexInMatrix(double d1, double d2, gMatrix *M, int i, int j, int k, int l)
{
  if(d1) {
    ok(gdp_matrix_add_double(M->M, i, j, d1), "LinAlg_AddComplexInMatrix");
    ok(gdp_matrix_add_double(M->M, k, l, d1), "LinAlg_AddComplexInMatrix");
  }
  if(d2) {
    ok(gdp_matrix_add_double(M->M, i, l, -d2), "LinAlg_AddComplexInMatrix");
    ok(gdp_matrix_add_double(M->M, k, j, d2), "LinAlg_AddComplexInMatrix");
  }
}

void LinAlg_AssembleMatrix(gMatrix *M) { ok(gdp_matrix_assemble(M->M), "LinAlg_AssembleMatrix"); }

void LinAlg_AssembleVector(gVector *V)
{
  to_device(V);
  ok(gdp_vector_assemble(V->V), "LinAlg_AssembleVector");
}

// ---- vector and matrix arithmetic (the library enforces the reference's
// "Wrong arguments" aliasing rules) -------------------------------------------------------------------

void LinAlg_AddVectorVector(gVector *V1, gVector *V2, gVector *V3)
{
  to_device(V2);
  device_write(V1);
  ok(gdp_vector_add(V1->V, V2->V, V3->V), "LinAlg_AddVectorVector");
}

void LinAlg_AddVectorProdVectorDouble(gVector *V1, gVector *V2, double d, gVector *V3)
{
  to_device(V2);
  device_write(V1);
  ok(gdp_vector_add_scaled(V1->V, V2->V, d, V3->V), "LinAlg_AddVectorProdVectorDouble");
}

void LinAlg_AddProdVectorDoubleProdVectorDouble(double alpha, gVector *V1, double beta, gVector *V2, gVector *V3)
{
  Message::Error("'LinAlg_AddProdVectorDoubleProdVectorDouble' not yet implemented");
}

void LinAlg_AddMatrixMatrix(gMatrix *M1, gMatrix *M2, gMatrix *M3)
{
  ok(gdp_matrix_add(M1->M, M2->M, M3->M), "LinAlg_AddMatrixMatrix");
}

void LinAlg_AddMatrixProdMatrixDouble(gMatrix *M1, gMatrix *M2, double d, gMatrix *M3)
{
  ok(gdp_matrix_add_scaled(M1->M, M2->M, d, M3->M), "LinAlg_AddMatrixProdMatrixDouble");
}

void LinAlg_SubVectorVector(gVector *V1, gVector *V2, gVector *V3)
{
  device_write(V1);
  device_write(V2);
  ok(gdp_vector_sub(V1->V, V2->V, V3->V), "LinAlg_SubVectorVector");
}

void LinAlg_SubMatrixMatrix(gMatrix *M1, gMatrix *M2, gMatrix *M3)
{
  Message::Error("'LinAlg_SubMatrixMatrix' not yet implemented");
}

void LinAlg_ProdVectorScalar(gVector *V1, gScalar *S, gVector *V2)
{
  Message::Error("'LinAlg_ProdVectorScalar' not yet implemented");
}

void LinAlg_ProdVectorDouble(gVector *V1, double d, gVector *V2)
{
  device_write(V1);
  ok(gdp_vector_scale(V1->V, d, V2->V), "LinAlg_ProdVectorDouble");
}

void LinAlg_ProdVectorComplex(gVector *V1, double d1, double d2, gVector *V2)
{
  Message::Error("'LinAlg_ProdVectorComplex' not yet implemented");
}

void LinAlg_ProdVectorVector(gVector *V1, gVector *V2, double *d)
{
  to_device(V1);
  to_device(V2);
  ok(gdp_vector_dot(V1->V, V2->V, d), "LinAlg_ProdVectorVector");
}

void LinAlg_ProdMatrixVector(gMatrix *M, gVector *V1, gVector *V2)
{
  if(V2 == V1) {
    Message::Error("Wrong arguments in 'LinAlg_ProdMatrixVector'");
    return;
  }
  to_device(V1);
  V2->mirror_state = 0;
  ok(gdp_prod_matrix_vector(M->M, V1->V, V2->V), "LinAlg_ProdMatrixVector");
}

void LinAlg_ProdMatrixScalar(gMatrix *M1, gScalar *S, gMatrix *M2)
{
  ok(gdp_matrix_scale(M1->M, S->s, M2->M), "LinAlg_ProdMatrixScalar");
}

void LinAlg_ProdMatrixDouble(gMatrix *M1, double d, gMatrix *M2)
{
  ok(gdp_matrix_scale(M1->M, d, M2->M), "LinAlg_ProdMatrixDouble");
}

void LinAlg_ProdMatrixComplex(gMatrix *M1, double d1, double d2, gMatrix *M2)
{
  Message::Error("'LinAlg_ProdMatrixComplex' not yet implemented");
}

void LinAlg_VectorNorm2(gVector *V1, double *norm)
{
  to_device(V1);
  ok(gdp_vector_norm2(V1->V, norm), "LinAlg_VectorNorm2");
}

void LinAlg_VectorNormInf(gVector *V1, double *norm)
{
  to_device(V1);
  ok(gdp_vector_norm_inf(V1->V, norm), "LinAlg_VectorNormInf");
}

// ---- solve -----------------------------------------------------------------------------------------

// The messages of solve_matrix (Sparskit.cpp:429-453, 556-589)
static void report(gSolver *Solver, int rc)
{
  gdp_solve_info info;
  memset(&info, 0, sizeof(info));
  gdp_solver_info(Solver->S, &info);
  if(rc == GDP_E_ZERO_PIVOT) {
    Message::Error("Zero pivot on line %d in ILU", info.ilu_row);
    return;
  }
  if(rc == GDP_E_ZERO_ROW) {
    Message::Error("%s", gdp_last_error());
    return;
  }
  if(!ok(rc, "LinAlg_Solve")) return;
  if(info.status == -1) Message::Warning("Iterative solver has iterated too many times");
  else if(info.status == -3) Message::Warning("Iterative solver is facing a break-down");
  else if(info.status < 0) Message::Warning("Iterative solver terminated (code = %d)", info.status);
  Message::Info("%d Iterations / Residual: %g", info.its, info.true_res);
}

void LinAlg_Solve(gMatrix *A, gVector *B, gSolver *Solver, gVector *X, int solverIndex)
{
  to_device(B);
  X->mirror_state = 0;
  report(Solver, gdp_solve(A->M, B->V, Solver->S, X->V));
}

void LinAlg_SolveAgain(gMatrix *A, gVector *B, gSolver *Solver, gVector *X, int solverIndex)
{
  to_device(B);
  X->mirror_state = 0;
  report(Solver, gdp_solve_again(A->M, B->V, Solver->S, X->V));
}

void LinAlg_SolveNL(gMatrix *A, gVector *B, gMatrix *Jac, gVector *R, gSolver *Solver, gVector *X, int solverIndex)
{
  Message::Error("'LinAlg_SolveNL' not yet implemented for Sparskit");
}

#endif
