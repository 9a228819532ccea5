// Solver parameters: the solver.par table of the reference's Sparskit backend
// (contrib/Sparskit/Sparskit.cpp:1847-1868 names and defaults, 1895-2000 file
// syntax, 2000-2042 per-option override) plus this backend's extra keys.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gdp_internal.h"

namespace gdp {

namespace {

enum Kind { INT, REAL };
struct Entry {
  const char *name;
  Kind kind;
  int Params::*ip;
  double Params::*dp;
};

const Entry kTable[] = {
  {"Matrix_Format", INT, &Params::Matrix_Format, nullptr},
  {"Matrix_Printing", INT, &Params::Matrix_Printing, nullptr},
  {"Matrix_Storage", INT, &Params::Matrix_Storage, nullptr},
  {"Scaling", INT, &Params::Scaling, nullptr},
  {"Renumbering_Technique", INT, &Params::Renumbering_Technique, nullptr},
  {"Preconditioner", INT, &Params::Preconditioner, nullptr},
  {"Preconditioner_Position", INT, &Params::Preconditioner_Position, nullptr},
  {"Nb_Fill", INT, &Params::Nb_Fill, nullptr},
  {"Permutation_Tolerance", REAL, nullptr, &Params::Permutation_Tolerance},
  {"Dropping_Tolerance", REAL, nullptr, &Params::Dropping_Tolerance},
  {"Diagonal_Compensation", REAL, nullptr, &Params::Diagonal_Compensation},
  {"Re_Use_ILU", INT, &Params::Re_Use_ILU, nullptr},
  {"Algorithm", INT, &Params::Algorithm, nullptr},
  {"Krylov_Size", INT, &Params::Krylov_Size, nullptr},
  {"IC_Acceleration", REAL, nullptr, &Params::IC_Acceleration},
  {"Re_Use_LU", INT, &Params::Re_Use_LU, nullptr},
  {"Iterative_Improvement", INT, &Params::Iterative_Improvement, nullptr},
  {"Nb_Iter_Max", INT, &Params::Nb_Iter_Max, nullptr},
  {"Stopping_Test", REAL, nullptr, &Params::Stopping_Test},
  // backend extensions
  {"Precision", INT, &Params::Precision, nullptr},
  {"ILU_Ordering", INT, &Params::ILU_Ordering, nullptr},
  {"Orthogonalization", INT, &Params::Orthogonalization, nullptr},
  {"ILU_Storage", INT, &Params::ILU_Storage, nullptr},
  {"Refinement_Steps", INT, &Params::Refinement_Steps, nullptr},
  {"Inner_Tolerance", REAL, nullptr, &Params::Inner_Tolerance},
  {"Inner_Krylov_Size", INT, &Params::Inner_Krylov_Size, nullptr},
};

const Entry *find(const char *name)
{
  for(const Entry &e : kTable)
    if(!std::strcmp(e.name, name)) return &e;
  return nullptr;
}

}  // namespace

// Per-option override.  Real values pass through a float, as in
// init_solver_option (Sparskit.cpp:2006-2014: `float valf = atof(value)`).
int params_set(Params &p, const char *name, const char *value)
{
  const Entry *e = find(name);
  if(!e) return set_error(GDP_E_ARG, "Unknown solver parameter '%s'", name);
  if(e->kind == INT) p.*(e->ip) = std::atoi(value);
  else {
    float f = (float)std::atof(value);
    p.*(e->dp) = (double)f;
  }
  return GDP_OK;
}

int params_get(const Params &p, const char *name, double *v)
{
  const Entry *e = find(name);
  if(!e) return set_error(GDP_E_ARG, "Unknown solver parameter '%s'", name);
  *v = e->kind == INT ? (double)(p.*(e->ip)) : p.*(e->dp);
  return GDP_OK;
}

// solver.par syntax (Sparskit.cpp:1968-2000): whitespace-separated
// "Name value" pairs, C comments skipped; real values read with %lf.
int params_read_file(Params &p, const char *file)
{
  FILE *f = std::fopen(file, "r");
  if(!f) return GDP_OK;  // defaults (the reference would prompt interactively)
  char buf[256];
  while(std::fscanf(f, "%255s", buf) == 1) {
    if(buf[0] == '/' && buf[1] == '*') {
      if(std::strstr(buf + 2, "*/")) continue;
      int prev = 0, c;
      while((c = std::fgetc(f)) != EOF) {
        if(prev == '*' && c == '/') break;
        prev = c;
      }
      continue;
    }
    const Entry *e = find(buf);
    if(!e) {
      char skip[256];
      if(std::fscanf(f, "%255s", skip) != 1) break;
      continue;
    }
    if(e->kind == INT) {
      int v;
      if(std::fscanf(f, "%d", &v) == 1) p.*(e->ip) = v;
    }
    else {
      double v;
      if(std::fscanf(f, "%lf", &v) == 1) p.*(e->dp) = v;
    }
  }
  std::fclose(f);
  return GDP_OK;
}

}  // namespace gdp
