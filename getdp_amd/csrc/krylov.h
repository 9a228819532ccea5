// Krylov solvers on the device (CG, BCGSTAB, TFQMR, GMRES, FGMRES of
// contrib/Sparskit/iters.f).
#pragma once

#include "gdp_device.h"
#include "ilu.h"
#include "sell.h"

namespace gdp {

constexpr int kMaxKrylov = 64;

// Control block kept in device memory: the scalars of Sparskit's ipar/fpar
// arrays that the iteration needs, the Givens rotations and the Hessenberg
// matrix.  Only `done`, `status` and `its` are read back by the host.
struct KCtl {
  int done, status, its, cycle_end;
  int kfinal, ierr12, ipar11, pend_flag;
  int maxits, m, nb, pad;
  double rtol, fpar3, fpar4, fpar5, fpar6, fpar7, alpha, beta;
  double scal, nrm0, thr, pend, omega;
  double tao, te, tao1, te1;  // TFQMR: tau and theta*eta after each half step
  double rr[2];  // CG: <r,r>, BCGSTAB: rho = <r,r0> of consecutive iterations (parity slots)
  double g[kMaxKrylov + 2], c[kMaxKrylov + 1], s[kMaxKrylov + 1];
  double h1[kMaxKrylov + 2], h2[kMaxKrylov + 2];
  double H[(kMaxKrylov + 1) * (kMaxKrylov + 2)];  // column j: H[j*(kMaxKrylov+2) + i]
};

// Operator: y = A x with values of type TA (SELL-64 layout, sell.h).
template <typename TA> using DevOp = SellOp<TA>;

// Preconditioner applied as z = M^{-1} v (right preconditioning).
enum PrecKind { PREC_NONE = 0, PREC_ILU = 1, PREC_USER = 2, PREC_INNER32 = 3 };
struct InnerSolver;  // FP32 GMRES used as the preconditioner of an FP64 solve
struct UserPrec;  // learned preconditioner (unet.hip)

// z = M^{-1} v for the learned preconditioner; skipped when the control block
// says so (gmode as for the other kernels of a cycle).
int user_prec_apply(UserPrec *u, const double *in, double *out, KCtl *C, int gmode);
int user_prec_apply(UserPrec *u, const float *in, float *out, KCtl *C, int gmode);
int user_prec_lsq(const UserPrec *u);  // two-stage with least-squares scaling
// the solver's renumbering (original -> internal, internal -> original) and
// diagonal scaling (by original index, may be null) for the next applications
void user_prec_bind(UserPrec *u, const int *perm, const int *rperm, const double *rs);

template <typename TV> struct PrecOp {
  PrecKind kind = PREC_NONE;
  IluFactor *ilu = nullptr;
  UserPrec *user = nullptr;
  InnerSolver *inner = nullptr;
};

// Mixed precision by nested Krylov: z = M^{-1} v is k steps of GMRES in
// FP32 (FP32 SELL values, FP32 vectors, ILU(0) with FP32 factors, CGS2) on
// A z = v from z = 0, inside an FP64 flexible GMRES whose residuals,
// Gram-Schmidt and updates stay in FP64.  Device-side only, so the outer
// restart cycle can still be captured in one graph.
struct InnerSolver {
  DevOp<float> A;
  PrecOp<float> M;
  int k = 20, n = 0;
  KCtl *C = nullptr;
  double *part = nullptr;
  float *vec = nullptr;  // (k + 4) vectors of n
  ~InnerSolver();
  int prepare();
};
int inner_apply(InnerSolver *s, const double *v, double *z, KCtl *outer, int gmode);

struct KrylovOptions {
  int algorithm = 8;  // 1 CG, 5 BCGSTAB, 6 TFQMR, 8 GMRES, 9 FGMRES
  int m = 30;
  int maxits = 1000;
  double rtol = 1e-10;
  int orth = 0;  // 0 MGS + selective reorthogonalisation, 1 CGS2
  bool use_graphs = true;
};

struct KrylovResult {
  int its = 0, status = 0;
  double res0 = 0, res_est = 0;
};

// Solve A sol = rhs from sol = 0 on the library stream.
template <typename TA, typename TV>
int krylov_solve(const DevOp<TA> &A, const PrecOp<TV> &M, const TV *rhs, TV *sol,
                 const KrylovOptions &opt, KrylovResult &res);

}  // namespace gdp
