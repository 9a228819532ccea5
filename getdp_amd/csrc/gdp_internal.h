// Internal declarations shared by the host and device translation units of
// libgdp_hip.  Not part of the C-ABI (see include/gdp_hip.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gdp_hip.h"

namespace gdp {

// ---- error reporting -------------------------------------------------------
int set_error(int code, const char *fmt, ...);

// ---- host pre-processing (host_prep.cpp) ------------------------------------
struct HostCSR {
  int n = 0;
  std::vector<int> ia, ja;
};

// Entity -> incident (element, local index) list, elements ascending.
struct Incidence {
  std::vector<int> ptr;   // nent + 2 (entities 1..nent)
  std::vector<int> item;  // element * 8 + local index
};

int number_edges(int nelem, const int *etype, const int *conn, int nvisit,
                 const int *visit, int *edges_out, int *edge_nodes, int max_edges);
// The same numbering computed on the device (edges.hip); used once the
// library is initialised unless GDP_HOST_PREP is set.
bool edges_on_device();
int number_edges_device(int nelem, const int *etype, const int *conn, int nvisit,
                        const int *visit, int *edges_out, int *edge_nodes, int max_edges);
int rcmk(int n, const int *ia, const int *ja, std::vector<int> &perm,
         std::vector<int> &rperm);
// Dependency depth (levels) of the strictly-lower part of a CSR matrix with
// sorted columns; lev[i] = level of row i (0-based).  Returns #levels.
int lower_levels(int n, const int *ia, const int *ja, std::vector<int> &lev);
int upper_levels(int n, const int *ia, const int *ja, std::vector<int> &lev);
// Greedy distance-1 colouring in natural order; returns #colours.
int greedy_colouring(int n, const int *ia, const int *ja, std::vector<int> &col);
// B = P A P^T with sorted columns: pattern of B plus map[kB] = kA.
void permute_pattern(int n, const int *ia, const int *ja, const int *perm,
                     std::vector<int> &ib, std::vector<int> &jb,
                     std::vector<int> &map);

// ---- ILUT / ILUTP factorisation (ilut.cpp) ------------------------------------
// Row i occupies [ptr[i], ptr[i+1]): its L entries, then at dpos[i] the
// inverted pivot (column i), then its U entries, each part in the order the
// triangular sweeps accumulate it; columns in the permuted numbering.
template <typename T> struct IlutRows {
  std::vector<int> ptr, dpos, col;
  std::vector<T> val;
  std::vector<int> newpos;  // ILUTP: permuted number of each column (empty: none)
};
template <typename T>
int ilut_factor(int n, const int *ia, const int *ja, const double *a, int lfil, double droptol,
                double permtol, bool pivot, IlutRows<T> &f);

// ---- solver parameters (solver_params.cpp) ---------------------------------
struct Params {
  int Matrix_Format = 1, Matrix_Printing = 0, Matrix_Storage = 0;
  int Scaling = 0, Renumbering_Technique = 1, Preconditioner = 2;
  int Preconditioner_Position = 2, Nb_Fill = 20;
  double Permutation_Tolerance = 5e-2, Dropping_Tolerance = 0.,
         Diagonal_Compensation = 0.;
  int Re_Use_ILU = 0, Algorithm = 8, Krylov_Size = 40;
  double IC_Acceleration = 1.;
  int Re_Use_LU = 0, Iterative_Improvement = 0, Nb_Iter_Max = 1000;
  double Stopping_Test = 1e-10;
  // backend extensions
  int Precision = 0, ILU_Ordering = 0, Orthogonalization = 0, ILU_Storage = 0;
  int Refinement_Steps = 20;
  double Inner_Tolerance = 1e-6;
  int Inner_Krylov_Size = 20;  // Precision 2: FP32 GMRES steps per preconditioner application
};
int params_set(Params &p, const char *name, const char *value);
int params_get(const Params &p, const char *name, double *v);
int params_read_file(Params &p, const char *file);

}  // namespace gdp
