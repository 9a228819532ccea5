// The MI355X branch of GetDP's backend types.  GetDP chooses its linear
// algebra backend at compile time (src/kernel/LinAlg.h:20-88); a maintainer
// adds one more branch in front of the Sparskit one:
//
//   #if defined(HAVE_GDP_HIP)
//   #include "LinAlg_GDPHIP.h"
//   #elif defined(HAVE_SPARSKIT)
//   ...
//
// and compiles LinAlg_GDPHIP.cpp (every LinAlg_* function of LinAlg.h:90-213)
// and Generate_GDPHIP.cpp (the device Generate[A] hook) in place of
// LinAlg_SPARSKIT.cpp.  Real arithmetic, like the Sparskit backend.
#ifndef LINALG_GDPHIP_H
#define LINALG_GDPHIP_H

#include "gdp_hip.h"

#define gSCALAR_SIZE 1
#define gCOMPLEX_INCREMENT 2

typedef struct {
  double s;
} gScalar;

// A device vector plus a host mirror for the entry-wise accessors
// (post-processing reads single entries, LinAlg_GetDoubleInVector): the
// mirror is refreshed from the device when stale, and entry-wise writes
// (LinAlg_SetDoubleInVector) go to the mirror, which is sent to the device
// before the next device operation.
typedef struct {
  int N;
  gdp_vector *V;
  double *mirror;
  int mirror_state;  // 0 stale, 1 equal to the device, 2 newer than the device
} gVector;

typedef struct {
  gdp_matrix *M;
} gMatrix;

typedef struct {
  gdp_solver *S;
} gSolver;

// DummyDof table of the current DofData (Current.DofData->DummyDof), provided
// by Generate_GDPHIP.cpp, which sees GetDP's global state.
int *GDP_CurrentDummyDof(void);

// Generate[A] of one formulation on the device (Generate_GDPHIP.cpp): 1 when
// done, 0 when GetDP's own element loop (Treatment_Formulation) must run.
struct Formulation;
struct DofData;
int GDP_GenerateFormulation(struct Formulation *Formulation_P, struct DofData *DofData_P);

#endif
