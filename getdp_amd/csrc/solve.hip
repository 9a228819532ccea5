// LinAlg_Solve / LinAlg_SolveAgain on the device: the control flow of the
// reference's solve_matrix (contrib/Sparskit/Sparskit.cpp:108-634):
//   diagonal scaling (Scaling 1, :211-218 + scale_matrix :962-1060)
//   zero-row check (:220-223)
//   RCMK renumbering (:231-259, cmkreord + sortcol)
//   ILU(0) factorisation (:417-452), FP32 factor storage by default
//   RHS permutation, Krylov iterations from x0 = 0 (:469-573)
//   true residual |A sol - rhs| (:578-589), inverse permutation (:598-602),
//   unscaling (:631-632)
// plus this backend's options: multicolour ILU ordering and mixed precision
// (FP32 inner Krylov solve with FP64 iterative refinement).
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <utility>

#include "krylov.h"
#include "sell.h"

namespace gdp {

namespace {

constexpr int BT = 256;

// row_of[k] = the row of entry k
__global__ void __launch_bounds__(BT) k_row_of(int n, const int *__restrict__ ia, int *__restrict__ row_of)
{
  const int i = blockIdx.x * BT + threadIdx.x;
  if(i < n)
    for(int k = ia[i]; k < ia[i + 1]; k++) row_of[k] = i;
}

__global__ void __launch_bounds__(BT) k_gather_scaled(long nnz, const int *__restrict__ map,
                                                     const int *__restrict__ row_of,
                                                     const int *__restrict__ ja_new,
                                                     const int *__restrict__ rperm,
                                                     const double *__restrict__ a_old,
                                                     const double *__restrict__ rs,
                                                     double *__restrict__ a_new)
{
  long k = (long)blockIdx.x * BT + threadIdx.x;
  if(k >= nnz) return;
  double v = a_old[map[k]];
  if(rs) {
    // diamua then amudia (blassm.f:785-870): a*rowscal(i), then *colscal(j)
    v = v * rs[rperm[row_of[k]]];
    v = v * rs[rperm[ja_new[k]]];
  }
  a_new[k] = v;
}

__global__ void __launch_bounds__(BT) k_diag_values(int n, const int *ia, const int *ja,
                                                   const double *a, double *d)
{
  int i = blockIdx.x * BT + threadIdx.x;
  if(i >= n) return;
  double v = 0.;
  for(int k = ia[i]; k < ia[i + 1]; k++)
    if(ja[k] == i) v = a[k];
  d[i] = v;
}

__global__ void __launch_bounds__(BT) k_permute_rhs(int n, const double *b, const double *rs,
                                                   const int *rperm, double *rhs)
{
  int i = blockIdx.x * BT + threadIdx.x;
  if(i >= n) return;
  int o = rperm[i];
  double v = b[o];
  if(rs) v *= rs[o];  // scale_vector(ROW) before the permutation
  rhs[i] = v;
}

__global__ void __launch_bounds__(BT) k_unpermute_sol(int n, const double *sol, const double *rs,
                                                     const int *perm, double *x)
{
  int i = blockIdx.x * BT + threadIdx.x;
  if(i >= n) return;
  double v = sol[perm[i]];
  if(rs) v *= rs[i];  // scale_vector(COLUMN)
  x[i] = v;
}

__global__ void __launch_bounds__(BT) k_to_f32(long n, const double *a, float *b)
{
  long i = (long)blockIdx.x * BT + threadIdx.x;
  if(i < n) b[i] = (float)a[i];
}

// r = rhs - A x (FP64, SELL layout)
__global__ void __launch_bounds__(256) k_residual(SellOp<double> A, const double *__restrict__ x,
                                                  const double *__restrict__ rhs,
                                                  double *__restrict__ r)
{
  const int row = blockIdx.x * 256 + threadIdx.x;
  if(row >= A.n) return;
  double sum;
  sell_row<double, double>(row, A, x, sum);
  r[row] = rhs[row] - sum;
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_sumsq(int n, const double *x, double *part)
{
  double s = 0;
  for(long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK)
    s += x[i] * x[i];
  s = block_sum<double, BLOCK>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// mixed precision: d32 = r / rnorm (FP32), later x += rnorm * d32
__global__ void __launch_bounds__(BT) k_scale_to_f32(int n, const double *r, double inv,
                                                    float *d)
{
  int i = blockIdx.x * BT + threadIdx.x;
  if(i < n) d[i] = (float)(r[i] * inv);
}
__global__ void __launch_bounds__(BT) k_add_scaled_f32(int n, double s, const float *d, double *x)
{
  int i = blockIdx.x * BT + threadIdx.x;
  if(i < n) x[i] += s * (double)d[i];
}

}  // namespace

int norm2_dev(const double *x, int n, double *out)
{
  const int nb = 512;
  static double *part = nullptr;
  if(!part) GDP_TRY(dalloc(&part, nb));
  hipLaunchKernelGGL(k_sumsq<256>, dim3(nb), dim3(256), 0, ctx().stream, n, x, part);
  double h[nb];
  GDP_TRY(d2h(h, part, nb));
  double s = 0;
  for(int i = 0; i < nb; i++) s += h[i];
  *out = std::sqrt(s);
  return GDP_OK;
}

int residual_dev(const DevSell &A, const double *x, const double *rhs, double *r)
{
  if(!A.n) return GDP_OK;
  hipLaunchKernelGGL(k_residual, dim3((A.n + 255) / 256), dim3(256), 0, ctx().stream, sell_op<double>(A), x,
                     rhs, r);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

// Renumbered / scaled copy of a matrix plus its factorisation, kept in the
// solver between LinAlg_Solve and LinAlg_SolveAgain.
struct SolveCache {
  const gdp_matrix *A = nullptr;
  uint64_t pattern_version = 0, value_version = 0;
  int renumbering = -1, ordering = -1, scaling = -1, precond = -1, ilu_f64 = -1;
  int n = 0;
  long nnz = 0;
  std::vector<int> perm, rperm, ib, jb, map;
  int *d_perm = nullptr, *d_rperm = nullptr, *d_map = nullptr, *d_row_of = nullptr;
  DevCSR Bm;              // renumbered (scaled) matrix
  DevSell sell;           // its SpMV layout (FP64 values, FP32 copy on demand)
  float *b32 = nullptr;   // FP32 copy of Bm's values
  double *rs = nullptr;   // scaling factors (null when not scaling)
  double *rhs = nullptr, *sol = nullptr, *work = nullptr;
  float *rhs32 = nullptr, *sol32 = nullptr;
  IluFactor ilu;   // ILU(0) on Bm's pattern
  IluFactor ilut;  // ILUT / ILUTP factors (own pattern)
  int ilut_fill = -1, ilut_storage = -1;
  double ilut_drop = -1., ilut_perm = -1.;
  std::unique_ptr<InnerSolver> inner;  // Precision 2
  int ilu_row = 0;
  bool have_factor = false;

  // the factors Preconditioner `prec` applies (null: none)
  IluFactor *factors(int prec)
  {
    if(prec == 6) return &ilu;
    if(prec == 1 || prec == 2) return &ilut;
    return nullptr;
  }

  void release()
  {
    dfree(d_perm);
    dfree(d_rperm);
    dfree(d_map);
    dfree(d_row_of);
    Bm.release();
    sell.release();
    dfree(b32);
    dfree(rs);
    dfree(rhs);
    dfree(sol);
    dfree(work);
    dfree(rhs32);
    dfree(sol32);
    ilu.release();
    ilut.release();
    ilut_fill = -1;
    have_factor = false;
    A = nullptr;
  }
  ~SolveCache() { release(); }
};

}  // namespace gdp

using namespace gdp;

struct gdp_solver {
  Params p;
  gdp_solve_info info{};
  SolveCache cache;
  UserPrec *user = nullptr;  // learned preconditioner (optional)
};

namespace gdp {
UserPrec *solver_user_prec(gdp_solver *s) { return s->user; }
void solver_set_user_prec(gdp_solver *s, UserPrec *u) { s->user = u; }
}

namespace {

double now_ms()
{
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Build (or reuse) the renumbered/scaled matrix and its ILU(0) factors.
int prepare(gdp_solver *S, const gdp_matrix *A, bool reuse_factor)
{
  SolveCache &c = S->cache;
  const Params &p = S->p;
  int prec = p.Preconditioner;
  bool same_pattern = c.A == A && c.pattern_version == A->pattern_version &&
                      c.renumbering == p.Renumbering_Technique && c.ordering == p.ILU_Ordering;
  bool same_values = same_pattern && c.value_version == A->value_version &&
                     c.scaling == p.Scaling;
  const bool same_ilut = (prec != 1 && prec != 2) ||
                         (c.ilut_fill == p.Nb_Fill && c.ilut_drop == p.Dropping_Tolerance &&
                          c.ilut_perm == p.Permutation_Tolerance && c.ilut_storage == p.ILU_Storage);
  if(reuse_factor && same_values && c.precond == prec && c.have_factor && same_ilut) return GDP_OK;
  const int n = A->n;
  const long nnz = A->csr.nnz;
  hipStream_t s = ctx().stream;
  if(!same_pattern) {
    c.release();
    c.n = n;
    c.nnz = nnz;
    // zero-row check (Sparskit.cpp:220-223)
    for(int i = 0; i < n; i++)
      if(A->h_ia[i + 1] - A->h_ia[i] <= 0) return set_error(GDP_E_ZERO_ROW, "Zero row in matrix (row %d)", i + 1);
    std::vector<int> perm(n), rperm(n);
    for(int i = 0; i < n; i++) perm[i] = rperm[i] = i;
    if(p.Renumbering_Technique == 1) rcmk(n, A->h_ia.data(), A->h_ja.data(), perm, rperm);
    else if(p.Renumbering_Technique != 0)
      return set_error(GDP_E_UNSUPPORTED, "Unknown renumbering technique");
    std::vector<int> ib, jb, map;
    permute_pattern(n, A->h_ia.data(), A->h_ja.data(), perm.data(), ib, jb, map);
    if(p.ILU_Ordering == 1) {
      // multicolour ordering on top: colour classes in order, rows of a class
      // keep their relative order -> the L factor has #colours levels
      std::vector<int> col;
      int ncol = greedy_colouring(n, ib.data(), jb.data(), col);
      std::vector<int> cnt(ncol + 1, 0), p2(n);
      for(int i = 0; i < n; i++) cnt[col[i] + 1]++;
      for(int k = 0; k < ncol; k++) cnt[k + 1] += cnt[k];
      for(int i = 0; i < n; i++) p2[i] = cnt[col[i]]++;
      std::vector<int> perm2(n), rperm2(n);
      for(int i = 0; i < n; i++) perm2[i] = p2[perm[i]];
      for(int i = 0; i < n; i++) rperm2[perm2[i]] = i;
      perm.swap(perm2);
      rperm.swap(rperm2);
      permute_pattern(n, A->h_ia.data(), A->h_ja.data(), perm.data(), ib, jb, map);
    }
    c.perm = std::move(perm);
    c.rperm = std::move(rperm);
    c.ib = std::move(ib);
    c.jb = std::move(jb);
    c.map = std::move(map);
    GDP_TRY(dalloc(&c.d_perm, (size_t)std::max(1, n)));
    GDP_TRY(dalloc(&c.d_rperm, (size_t)std::max(1, n)));
    GDP_TRY(dalloc(&c.d_map, (size_t)std::max(1L, nnz)));
    GDP_TRY(dalloc(&c.d_row_of, (size_t)std::max(1L, nnz)));
    GDP_TRY(h2d(c.d_perm, c.perm.data(), (size_t)n));
    GDP_TRY(h2d(c.d_rperm, c.rperm.data(), (size_t)n));
    GDP_TRY(h2d(c.d_map, c.map.data(), (size_t)nnz));
    c.Bm.n = n;
    c.Bm.nnz = nnz;
    GDP_TRY(dalloc(&c.Bm.ia, (size_t)n + 1));
    GDP_TRY(dalloc(&c.Bm.ja, (size_t)std::max(1L, nnz)));
    GDP_TRY(dalloc(&c.Bm.a, (size_t)std::max(1L, nnz)));
    GDP_TRY(h2d(c.Bm.ia, c.ib.data(), (size_t)n + 1));
    GDP_TRY(h2d(c.Bm.ja, c.jb.data(), (size_t)nnz));
    if(n) hipLaunchKernelGGL(k_row_of, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, c.Bm.ia, c.d_row_of);
    GDP_HIP(hipGetLastError());
    GDP_TRY(dalloc(&c.rhs, (size_t)std::max(1, n)));
    GDP_TRY(dalloc(&c.sol, (size_t)std::max(1, n)));
    GDP_TRY(dalloc(&c.work, (size_t)std::max(1, n)));
    c.A = A;
    c.pattern_version = A->pattern_version;
    c.renumbering = p.Renumbering_Technique;
    c.ordering = p.ILU_Ordering;
    c.value_version = 0;
  }
  // values: scaling, permutation
  if(!same_values) {
    dfree(c.rs);
    if(p.Scaling == 1) {
      double *d = nullptr;
      GDP_TRY(dalloc(&d, (size_t)std::max(1, n)));
      if(n) hipLaunchKernelGGL(k_diag_values, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, A->csr.ia, A->csr.ja, A->csr.a, d);
      std::vector<double> h(n);
      GDP_TRY(d2h(h.data(), d, (size_t)n));
      dfree(d);
      bool ok = true;
      for(int i = 0; i < n; i++) {
        if(h[i]) h[i] = 1. / std::sqrt(std::fabs(h[i]));
        else {
          ok = false;  // "Diagonal scaling aborted because of zero diagonal element"
          break;
        }
      }
      if(ok) {
        GDP_TRY(dalloc(&c.rs, (size_t)std::max(1, n)));
        GDP_TRY(h2d(c.rs, h.data(), (size_t)n));
      }
    }
    else if(p.Scaling != 0)
      return set_error(GDP_E_UNSUPPORTED, "Scaling %d not supported (only 0 and 1)", p.Scaling);
    if(nnz)
      hipLaunchKernelGGL(k_gather_scaled, dim3((unsigned)((nnz + BT - 1) / BT)), dim3(BT), 0, s, nnz,
                         c.d_map, c.d_row_of, c.Bm.ja, c.d_rperm, A->csr.a, c.rs, c.Bm.a);
    GDP_HIP(hipGetLastError());
    if(!c.sell.n) GDP_TRY(sell_build(c.ib, c.Bm, c.sell));
    else GDP_TRY(sell_copy_values(c.Bm, c.sell, false));
    dfree(c.sell.val32);
    c.value_version = A->value_version;
    c.scaling = p.Scaling;
    c.have_factor = false;
  }
  if(p.Precision >= 1 && !c.sell.val32) GDP_TRY(sell_copy_values(c.Bm, c.sell, true));
  // factorisation
  c.precond = prec;
  if(prec == 6) {
    GDP_TRY(sync());
    const double ta = now_ms();
    if(!c.ilu.diag) GDP_TRY(c.ilu.analyse(c.Bm, c.ib, c.jb));  // first ILU(0) on this pattern
    GDP_TRY(sync());
    const double tb = now_ms();
    GDP_TRY(c.ilu.factor(c.Bm.a, p.ILU_Storage == 1, &c.ilu_row));
    GDP_TRY(sync());
    S->info.t_load_ms = tb - ta;
    S->info.t_factor_ms = now_ms() - tb;
    if(c.ilu_row) {
      c.have_factor = false;
      return set_error(GDP_E_ZERO_PIVOT, "Zero pivot on line %d in ILU", c.ilu_row);
    }
    c.have_factor = true;
  }
  else if(prec == 1 || prec == 2) {
    // ILUT / ILUTP (Sparskit.cpp:375-390) of the renumbered, scaled matrix,
    // factored on the host; the sweeps run on the device.  ILUTP's column
    // permutation stays inside the preconditioner: Krylov on A with
    // M^-1 = Q (LU)^-1, (Q v)[j] = v[newpos[j]], builds the same iterates as
    // the reference's Krylov on the column-permuted A Q with (LU)^-1, and its
    // solution is already in A's column numbering.
    c.have_factor = false;
    c.ilut_fill = -1;
    std::vector<double> hv(nnz);
    GDP_TRY(d2h(hv.data(), c.Bm.a, (size_t)nnz));
    const bool pivot = prec == 2;
    const double ta = now_ms();
    double tb = ta;
    if(p.ILU_Storage == 1) {
      IlutRows<double> f;
      GDP_TRY(ilut_factor(n, c.ib.data(), c.jb.data(), hv.data(), p.Nb_Fill, p.Dropping_Tolerance,
                          p.Permutation_Tolerance, pivot, f));
      tb = now_ms();
      GDP_TRY(c.ilut.load(f, true));
    }
    else {
      IlutRows<float> f;
      GDP_TRY(ilut_factor(n, c.ib.data(), c.jb.data(), hv.data(), p.Nb_Fill, p.Dropping_Tolerance,
                          p.Permutation_Tolerance, pivot, f));
      tb = now_ms();
      GDP_TRY(c.ilut.load(f, false));
    }
    GDP_TRY(sync());
    S->info.t_factor_ms = tb - ta;
    S->info.t_load_ms = now_ms() - tb;
    c.ilut_fill = p.Nb_Fill;
    c.ilut_drop = p.Dropping_Tolerance;
    c.ilut_perm = p.Permutation_Tolerance;
    c.ilut_storage = p.ILU_Storage;
    c.have_factor = true;
  }
  else if(prec == 0 || prec == 8)
    c.have_factor = true;  // NONE, or DIAGONAL = identity (Sparskit.cpp:338-354)
  else
    return set_error(GDP_E_UNSUPPORTED, "Preconditioner %d not supported (0, 1, 2, 6, 8)", prec);
  return GDP_OK;
}

int do_solve(gdp_solver *S, gdp_matrix *A, gdp_vector *b, gdp_vector *x, bool again)
{
  if(!S || !A || !b || !x) return set_error(GDP_E_ARG, "gdp_solve: null argument");
  if(A->n != b->n || A->n != x->n) return set_error(GDP_E_ARG, "gdp_solve: size mismatch");
  if(!A->csr.ia) return set_error(GDP_E_STATE, "gdp_solve: matrix has no assembled pattern");
  GDP_TRY(ensure_ready());
  GDP_TRY(vec_flush(b));
  if(x != b) x->staged.clear();  // the solve starts from x = 0 and overwrites x
  const Params &p = S->p;
  gdp_solve_info &info = S->info;
  std::memset(&info, 0, sizeof(info));
  const int n = A->n;
  if(n == 0) return GDP_OK;  // "No equations in linear system"
  hipStream_t s = ctx().stream;
  GDP_TRY(sync());
  double t0 = now_ms();
  int r = prepare(S, A, again || p.Re_Use_ILU);
  if(r) {
    info.ilu_row = S->cache.ilu_row;
    return r;
  }
  SolveCache &c = S->cache;
  GDP_TRY(sync());
  double t1 = now_ms();
  info.t_setup_ms = t1 - t0;
  IluFactor *F = c.factors(p.Preconditioner);
  info.ilu_levels = F ? F->nlevels : 0;
  info.ilu_sweep = F ? F->sweep_kind() : 0;
  hipLaunchKernelGGL(k_permute_rhs, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, b->d, c.rs, c.d_rperm, c.rhs);
  GDP_HIP(hipGetLastError());
  double bnorm;
  GDP_TRY(norm2_dev(c.rhs, n, &bnorm));

  KrylovOptions opt;
  opt.algorithm = p.Algorithm;
  opt.m = p.Krylov_Size;
  opt.maxits = p.Nb_Iter_Max;
  opt.rtol = p.Stopping_Test;
  opt.orth = p.Orthogonalization;
  if(p.Preconditioner_Position != 2 && F)
    return set_error(GDP_E_UNSUPPORTED, "only right preconditioning (Preconditioner_Position 2)");
  KrylovResult kr;
  UserPrec *up = S->user;
  if(up) user_prec_bind(up, c.d_perm, c.d_rperm, c.rs);
  if(p.Precision == 2) {
    // FP64 flexible GMRES (residuals, Gram-Schmidt, updates) preconditioned
    // by Inner_Krylov_Size steps of FP32 GMRES (FP32 SpMV + FP32 ILU(0))
    DevOp<double> op = sell_op<double>(c.sell);
    if(!c.inner) c.inner.reset(new InnerSolver);
    InnerSolver &in = *c.inner;
    in.A = sell_op<float>(c.sell);
    in.M = PrecOp<float>();
    if(F) {
      in.M.kind = PREC_ILU;
      in.M.ilu = F;
    }
    if(in.k != p.Inner_Krylov_Size) {
      dfree(in.vec);
      in.k = p.Inner_Krylov_Size;
    }
    GDP_TRY(in.prepare());
    PrecOp<double> M;
    M.kind = PREC_INNER32;
    M.inner = &in;
    KrylovOptions o2 = opt;
    o2.algorithm = 9;
    GDP_TRY((krylov_solve<double, double>(op, M, c.rhs, c.sol, o2, kr)));
    if(F) GDP_TRY(F->check_timeout());
    info.its = kr.its;
    info.status = kr.status;
    info.res0 = kr.res0;
    info.res_est = kr.res_est;
    info.outer_its = kr.its;
  }
  else if(p.Precision == 0) {
    DevOp<double> op = sell_op<double>(c.sell);
    PrecOp<double> M;
    if(up && p.Algorithm == 9) {
      M.kind = PREC_USER;
      M.user = up;
      if(p.Preconditioner == 6) M.ilu = &c.ilu;  // network, then ILU(0) correction
    }
    else if(F) {
      M.kind = PREC_ILU;
      M.ilu = F;
    }
    GDP_TRY((krylov_solve<double, double>(op, M, c.rhs, c.sol, opt, kr)));
    info.its = kr.its;
    info.status = kr.status;
    info.res0 = kr.res0;
    info.res_est = kr.res_est;
  }
  else {
    // FP32 inner solves, FP64 outer residual and update (iterative refinement)
    if(!c.rhs32) {
      GDP_TRY(dalloc(&c.rhs32, (size_t)n));
      GDP_TRY(dalloc(&c.sol32, (size_t)n));
    }
    DevOp<float> op = sell_op<float>(c.sell);
    PrecOp<float> M;
    if(up && p.Algorithm == 9) {
      M.kind = PREC_USER;
      M.user = up;
      if(p.Preconditioner == 6) M.ilu = &c.ilu;
    }
    else if(F) {
      M.kind = PREC_ILU;
      M.ilu = F;
    }
    KrylovOptions in = opt;
    in.rtol = p.Inner_Tolerance;
    GDP_HIP(hipMemsetAsync(c.sol, 0, sizeof(double) * n, s));
    double target = p.Stopping_Test * bnorm;
    int total = 0, outer = 0, last_status = 0;
    double rn = bnorm;
    info.res0 = bnorm;
    for(outer = 0; outer <= p.Refinement_Steps; outer++) {
      GDP_TRY(residual_dev(c.sell, c.sol, c.rhs, c.work));
      GDP_TRY(norm2_dev(c.work, n, &rn));
      if(rn <= target || outer == p.Refinement_Steps) break;
      hipLaunchKernelGGL(k_scale_to_f32, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, c.work, 1.0 / rn, c.rhs32);
      in.maxits = p.Nb_Iter_Max > 0 ? std::max(1, p.Nb_Iter_Max - total) : 0;
      GDP_TRY((krylov_solve<float, float>(op, M, c.rhs32, c.sol32, in, kr)));
      total += kr.its;
      last_status = kr.status;
      hipLaunchKernelGGL(k_add_scaled_f32, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, rn, c.sol32, c.sol);
      GDP_HIP(hipGetLastError());
      if(p.Nb_Iter_Max > 0 && total >= p.Nb_Iter_Max) {
        GDP_TRY(residual_dev(c.sell, c.sol, c.rhs, c.work));
        GDP_TRY(norm2_dev(c.work, n, &rn));
        break;
      }
    }
    info.its = total;
    info.outer_its = outer;
    info.res_est = rn;
    info.status = rn <= target ? 0 : (last_status < 0 ? last_status : -1);
  }
  GDP_TRY(sync());
  double t2 = now_ms();
  info.t_solve_ms = t2 - t1;
  // true residual of the solved (renumbered, scaled) system
  GDP_TRY(residual_dev(c.sell, c.sol, c.rhs, c.work));
  GDP_TRY(norm2_dev(c.work, n, &info.true_res));
  info.rel_res = bnorm > 0 ? info.true_res / bnorm : info.true_res;
  hipLaunchKernelGGL(k_unpermute_sol, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, c.sol, c.rs, c.d_perm, x->d);
  GDP_HIP(hipGetLastError());
  GDP_TRY(sync());
  if(!p.Re_Use_ILU && !again) {
    // the reference frees the factors after each solve unless Re_Use_ILU
    // (Sparskit.cpp:606-623); they are kept here for LinAlg_SolveAgain.
  }
  return GDP_OK;
}

}  // namespace

extern "C" int gdp_solver_create(const char *file, gdp_solver **out)
{
  if(!out) return set_error(GDP_E_ARG, "gdp_solver_create: null output");
  auto *S = new gdp_solver;
  if(file) {
    int r = params_read_file(S->p, file);
    if(r) {
      delete S;
      return r;
    }
  }
  *out = S;
  return GDP_OK;
}

extern "C" int gdp_solver_set(gdp_solver *s, const char *name, const char *value)
{
  if(!s || !name || !value) return set_error(GDP_E_ARG, "gdp_solver_set: null argument");
  return params_set(s->p, name, value);
}

extern "C" int gdp_solver_get(const gdp_solver *s, const char *name, double *value)
{
  if(!s || !name || !value) return set_error(GDP_E_ARG, "gdp_solver_get: null argument");
  return params_get(s->p, name, value);
}

extern "C" int gdp_solver_destroy(gdp_solver *s)
{
  if(!s) return GDP_OK;
  if(ctx().ready) (void)sync();
  delete s;
  return GDP_OK;
}

extern "C" int gdp_solver_info(const gdp_solver *s, gdp_solve_info *info)
{
  if(!s || !info) return set_error(GDP_E_ARG, "gdp_solver_info: null argument");
  *info = s->info;
  return GDP_OK;
}

extern "C" int gdp_solve(gdp_matrix *A, gdp_vector *b, gdp_solver *s, gdp_vector *x)
{
  return do_solve(s, A, b, x, false);
}

extern "C" int gdp_solve_again(gdp_matrix *A, gdp_vector *b, gdp_solver *s, gdp_vector *x)
{
  return do_solve(s, A, b, x, true);
}

extern "C" int gdp_solver_bench_spmv(gdp_solver *s, int reps, double *avg_ms, int *compressed)
{
  if(!s || reps < 1 || !avg_ms || !compressed) return set_error(GDP_E_ARG, "gdp_solver_bench_spmv: bad arguments");
  const DevSell &S = s->cache.sell;
  if(!S.n) return set_error(GDP_E_STATE, "the solver has not solved a system yet");
  double *x = nullptr, *y = nullptr;
  GDP_TRY(dalloc(&x, (size_t)S.n));
  int rc = dalloc(&y, (size_t)S.n);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  if(!rc) rc = hipMemsetAsync(x, 0, sizeof(double) * S.n, ctx().stream) == hipSuccess ? GDP_OK
                                                                                     : set_error(GDP_E_HIP, "hipMemsetAsync");
  if(!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = set_error(GDP_E_HIP, "hipEventCreate");
  if(!rc) rc = sync();
  if(!rc && hipEventRecord(e0, ctx().stream) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventRecord");
  for(int r = 0; !rc && r < reps; r++) rc = sell_spmv(S, x, y);
  if(!rc && hipEventRecord(e1, ctx().stream) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventRecord");
  if(!rc && hipEventSynchronize(e1) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventSynchronize");
  if(!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventElapsedTime");
  if(e0) (void)hipEventDestroy(e0);
  if(e1) (void)hipEventDestroy(e1);
  dfree(x);
  dfree(y);
  GDP_TRY(rc);
  *avg_ms = ms / reps;
  *compressed = S.c16 != nullptr;
  return GDP_OK;
}

// out = (LU)^{-1} in with the ILU(0) factors of the solver's last solve, in
// its internal numbering (renumbered, scaled system), `reps` times in a row;
// *avg_ms (optional) = average time per application (HIP events).
extern "C" int gdp_solver_prec_apply(gdp_solver *s, const gdp_vector *in, gdp_vector *out, int reps,
                                     double *avg_ms)
{
  if(!s || !in || !out || reps < 1) return set_error(GDP_E_ARG, "gdp_solver_prec_apply: bad arguments");
  SolveCache &c = s->cache;
  IluFactor *F = c.factors(c.precond);
  if(!c.have_factor || !F || !c.n)
    return set_error(GDP_E_STATE, "the solver holds no ILU factors (no solve with Preconditioner 1, 2 or 6)");
  if(in->n != c.n || out->n != c.n || in == out)
    return set_error(GDP_E_ARG, "gdp_solver_prec_apply: vectors must be distinct and of the system's size");
  GDP_TRY(vec_flush(in));
  out->staged.clear();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  int rc = sync();
  if(!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = set_error(GDP_E_HIP, "hipEventCreate");
  if(!rc && hipEventRecord(e0, ctx().stream) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventRecord");
  for(int r = 0; !rc && r < reps; r++) rc = F->apply<double>(in->d, 1.0, out->d);
  if(!rc && hipEventRecord(e1, ctx().stream) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventRecord");
  if(!rc && hipEventSynchronize(e1) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventSynchronize");
  if(!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = set_error(GDP_E_HIP, "hipEventElapsedTime");
  if(e0) (void)hipEventDestroy(e0);
  if(e1) (void)hipEventDestroy(e1);
  GDP_TRY(rc);
  GDP_TRY(F->check_timeout());
  if(avg_ms) *avg_ms = ms / reps;
  return GDP_OK;
}
