// Krylov solvers on the device, following contrib/Sparskit/iters.f:
//   gmres  (iters.f:1912-2217)  restarted GMRES, right preconditioning
//   fgmres (iters.f:2522-2807)  flexible GMRES (stores Z = M^{-1} V)
//   cg     (iters.f:379-562)    Sparskit's CG on A M^{-1} (not PCG)
//   bcgstab (iters.f:1032-1306) BiCGStab on A M^{-1}, step mapped back at the end
//   tfqmr  (iters.f:1307-1584)  transpose-free QMR on A M^{-1}, likewise
// with x0 = 0 and the stopping rule of ipar(3) = 1: stop when the (estimated)
// residual norm falls to Stopping_Test times the first residual norm.
//
// Every scalar decision (Givens rotations, residual estimate, restart,
// stopping, breakdown) is taken on the device, so a whole restart cycle (or a
// batch of CG iterations) is enqueued without host round trips; it is
// captured once into a hipGraph and replayed.  Kernels of a cycle that has
// already ended return immediately.
//
// Reductions: a streaming kernel writes one partial sum per block; a
// one-block finalize kernel sums the partials in a fixed order and applies
// the scalar logic.  The kernel boundary makes the partials visible (no
// device-scope fences, which cost an L2 write-back on a multi-die GPU), and
// the result does not depend on scheduling.  Dot products accumulate in FP64
// whatever the vector type.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "krylov.h"
#include "sell.h"

namespace gdp {

namespace {

constexpr int B = 256;  // threads per block for vector kernels
constexpr int LDH = kMaxKrylov + 2;
// entries per thread in the CGS kernels: 4 on large vectors, 1 below
// kCgsSmall entries so that a small system still launches >= 256 workgroups
constexpr int kCgsSmall = 256 * 256 * 4;
inline int cgs_per(int n) { return n >= kCgsSmall ? 4 : 1; }

// gate modes: 0 none, 1 skip when done, 2 skip when done or cycle ended
__device__ inline bool gated(const KCtl *C, int mode)
{
  if(mode == 0) return false;
  if(mode == 1) return C->done != 0;
  return (C->done | C->cycle_end) != 0;
}

// givens (iters.f:3235-3272)
__device__ inline void givens(double &x, double &y, double &c, double &s)
{
  if(x == 0.0 && y == 0.0) {
    c = 1.0;
    s = 0.0;
  }
  else if(fabs(y) > fabs(x)) {
    double t = x / y;
    x = sqrt(1.0 + t * t);
    s = copysign(1.0 / x, y);
    c = t * s;
  }
  else if(fabs(y) <= fabs(x)) {
    double t = y / x;
    y = sqrt(1.0 + t * t);
    c = copysign(1.0 / y, x);
    s = t * c;
  }
  else {
    x = 0.0;
    y = 0.0;
    c = 1.0;
    s = 0.0;
  }
  x = fabs(x * y);
}

// End of Arnoldi step k (gmres label 50, iters.f:2076-2109), one thread.
__device__ inline void gm_givens_step(int k, KCtl *C)
{
  C->its++;
  if(C->ierr12 < 0) {
    C->cycle_end = 1;
    C->kfinal = k;
    return;
  }
  double *hc = C->H + (long)(k - 1) * LDH;
  for(int i = 0; i < k - 1; i++) {
    double a0 = hc[i];
    hc[i] = C->c[i] * a0 + C->s[i] * hc[i + 1];
    hc[i + 1] = C->c[i] * hc[i + 1] - C->s[i] * a0;
  }
  double c, s;
  givens(hc[k - 1], hc[k], c, s);
  C->c[k - 1] = c;
  C->s[k - 1] = s;
  double alpha = -s * C->g[k - 1];
  C->g[k - 1] = c * C->g[k - 1];
  C->g[k] = alpha;
  alpha = fabs(alpha);
  C->fpar5 = alpha;
  if(!(k < C->m && !(alpha <= C->fpar4) && (C->maxits <= 0 || C->its < C->maxits))) {
    C->cycle_end = 1;
    C->kfinal = k;
  }
}


// Finalisations shared by the one-block kernels and by the last block of a
// fused producer: mgsro's nrm0 = |v|^2 with the abnormal-number checks, and
// the Hessenberg column of step k with nrm1 = |V_k| (MGS: h1, CGS2: h1 + h2).
__device__ inline void fin_nrm0(KCtl *C, double tot)
{
  C->nrm0 = tot;
  C->thr = tot * 0.98;
  C->pend_flag = 0;
  if(tot <= 0.0) C->ierr12 = -1;
  else if(!(tot > 0.0 && 1.0 / tot > 0.0)) C->ierr12 = -2;
}

__device__ inline void fin_hcol(KCtl *C, int k, double tot, bool cgs)
{
  double nrm1 = sqrt(tot);
  double *hc = C->H + (long)(k - 1) * LDH;
  for(int i = 0; i < k; i++) hc[i] = cgs ? C->h1[i] + C->h2[i] : C->h1[i];
  hc[k] = nrm1;
  C->pend_flag = 0;
  if(nrm1 <= 0.0) C->ierr12 = -3;
  else C->scal = 1.0 / nrm1;
}

// out[i] = sum of part[i * np + j] over j, i < k: one wave per i, each lane
// adds a strided subset in order, then the wave adds its lanes (a fixed order).
template <int BLOCK> __device__ inline void wave_sums(const double *part, int np, int k, double *out)
{
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  for(int i = wv; i < k; i += BLOCK / kWave) {
    const double *p = part + (long)i * np;
    double v = 0.;
    for(int j = lane; j < np; j += kWave) v += p[j];
    v = wave_sum(v);
    if(lane == 0) out[i] = v;
  }
}

// ---- operator -------------------------------------------------------------------
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) k_op(SellOp<TA> A, const TV *__restrict__ x,
                                            TV *__restrict__ y, const KCtl *C, int gmode)
{
  if(gated(C, gmode)) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if(row >= A.n) return;
  double sum;
  sell_row<TA, TV>(row, A, x, sum);
  y[row] = (TV)sum;
}

// y = A x fused with the partial sums of |y|^2 (nrm0 of mgsro) or of <d, y>.
template <typename TA, typename TV, bool WITH_D>
__global__ void __launch_bounds__(256) k_op_dot(SellOp<TA> A, const TV *__restrict__ x,
                                                const TV *__restrict__ d, TV *__restrict__ y,
                                                const KCtl *C, int gmode, double *part)
{
  if(gated(C, gmode)) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  double t = 0.;
  if(row < A.n) {
    double sum;
    sell_row<TA, TV>(row, A, x, sum);
    TV yv = (TV)sum;
    y[row] = yv;
    t = WITH_D ? (double)d[row] * (double)yv : (double)yv * (double)yv;
  }
  t = block_sum<double, 256>(t);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

// t = v - A x (gated): the residual the learned stage leaves, one pass.
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) k_op_resid(SellOp<TA> A, const TV *__restrict__ x,
                                                  const TV *__restrict__ v, TV *__restrict__ t,
                                                  const KCtl *C, int gmode)
{
  if(gated(C, gmode)) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  if(row >= A.n) return;
  double sum;
  sell_row<TA, TV>(row, A, x, sum);
  t[row] = v[row] - (TV)sum;
}

// Scale of the learned stage (two-stage preconditioner): w = A x into t, and
// per-block partial sums of <v, w> and <w, w> into part[2 b], part[2 b + 1].
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) k_op_dots(SellOp<TA> A, const TV *__restrict__ x,
                                                 const TV *__restrict__ v, TV *__restrict__ t, double *part,
                                                 const KCtl *C, int gmode)
{
  if(gated(C, gmode)) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  double vw = 0., ww = 0.;
  if(row < A.n) {
    double sum;
    sell_row<TA, TV>(row, A, x, sum);
    const TV w = (TV)sum;
    t[row] = w;
    vw = (double)v[row] * (double)w;
    ww = (double)w * (double)w;
  }
  vw = block_sum<double, 256>(vw);
  __syncthreads();
  ww = block_sum<double, 256>(ww);
  if(threadIdx.x == 0) {
    part[2 * blockIdx.x] = vw;
    part[2 * blockIdx.x + 1] = ww;
  }
}

// alpha = <v, w> / <w, w> (the least-squares scale of the learned correction,
// 1 when w = 0), summed over the blocks in order, into part[2 nb]
__global__ void __launch_bounds__(256) k_alpha(int nb, double *part, const KCtl *C, int gmode)
{
  if(gated(C, gmode)) return;
  double vw = 0., ww = 0.;
  for(int b = threadIdx.x; b < nb; b += 256) {
    vw += part[2 * b];
    ww += part[2 * b + 1];
  }
  vw = block_sum<double, 256>(vw);
  __syncthreads();
  ww = block_sum<double, 256>(ww);
  if(threadIdx.x == 0) part[2 * nb] = (ww > 0. && vw == vw && ww < 1e300) ? vw / ww : 1.;
}

// x *= alpha; t = v - alpha w (t holds w)
template <typename TV>
__global__ void __launch_bounds__(256) k_scale_resid(int n, const double *alpha_p, TV *__restrict__ x,
                                                     const TV *__restrict__ v, TV *__restrict__ t,
                                                     const KCtl *C, int gmode)
{
  if(gated(C, gmode)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i >= n) return;
  const TV a = (TV)*alpha_p;
  x[i] = a * x[i];
  t[i] = v[i] - a * t[i];
}

template <typename TA, typename TV>
int op_apply(const DevOp<TA> &A, const TV *x, TV *y, const KCtl *C, int gmode)
{
  if(!A.n) return GDP_OK;
  hipLaunchKernelGGL((k_op<TA, TV>), dim3((A.n + 255) / 256), dim3(256), 0, ctx().stream, A, x, y, C,
                     gmode);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

// ---- streaming kernels with per-block partial sums ------------------------------------

// w = rhs - w (A*sol already in w); partial |w|^2
template <typename TV>
__global__ void __launch_bounds__(B) k_gm_resid(int n, const TV *__restrict__ rhs, TV *w,
                                               const KCtl *C, double *part)
{
  if(C->done) return;
  double s = 0;
  for(long i = (long)blockIdx.x * B + threadIdx.x; i < n; i += (long)gridDim.x * B) {
    TV v = rhs[i] - w[i];
    w[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename TV>
__global__ void __launch_bounds__(B) k_scale(int n, TV *v, const KCtl *C, int gmode)
{
  if(gated(C, gmode)) return;
  double f = C->scal;
  for(long i = (long)blockIdx.x * B + threadIdx.x; i < n; i += (long)gridDim.x * B)
    v[i] = (TV)(f * (double)v[i]);
}

// MGS step i, part 1: apply a pending re-orthogonalisation against V_{i-1},
// then partial sums of V_k . V_i.
template <typename TV>
__global__ void __launch_bounds__(B) k_mgs_dot(int n, int i, TV *__restrict__ vk,
                                              const TV *__restrict__ V, const KCtl *C, double *part)
{
  if(gated(C, 2) || C->ierr12) return;
  const TV *vi = V + (long)i * n;
  const TV *vp = i > 0 ? V + (long)(i - 1) * n : nullptr;
  const bool pend = C->pend_flag != 0;
  const double pc = C->pend;
  double s = 0;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) {
    TV x = vk[j];
    if(pend) {
      x = (TV)((double)x - pc * (double)vp[j]);
      vk[j] = x;
    }
    s += (double)x * (double)vi[j];
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// MGS step i, part 2: V_k -= h_i V_i, then partial sums of V_k . V_i for the
// re-orthogonalisation test of mgsro (iters.f:3563-3575).
template <typename TV>
__global__ void __launch_bounds__(B) k_mgs_axpy(int n, int i, TV *__restrict__ vk,
                                               const TV *__restrict__ V, const KCtl *C, double *part)
{
  if(gated(C, 2) || C->ierr12) return;
  const TV *vi = V + (long)i * n;
  const double f = C->h1[i];
  double s = 0;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) {
    TV x = (TV)((double)vk[j] - f * (double)vi[j]);
    vk[j] = x;
    s += (double)x * (double)vi[j];
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// MGS end: apply the last pending correction; partial |V_k|^2.
template <typename TV>
__global__ void __launch_bounds__(B) k_mgs_norm(int n, int k, TV *__restrict__ vk,
                                               const TV *__restrict__ V, const KCtl *C, double *part)
{
  if(gated(C, 2) || C->ierr12) return;
  const TV *vp = V + (long)(k - 1) * n;
  const bool pend = C->pend_flag != 0;
  const double pc = C->pend;
  double s = 0;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) {
    TV x = vk[j];
    if(pend) {
      x = (TV)((double)x - pc * (double)vp[j]);
      vk[j] = x;
    }
    s += (double)x * (double)x;
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// CGS2: pass 1 h = V^T v; pass 2 v -= V h, h' = V^T v; pass 3 v -= V h', |v|^2.
// One block owns one chunk of CGS_PER*B entries (grid = ceil(n / chunk)).
// Pass 1 starts with mgsro's nrm0 test on the partials of |w|^2 left by the
// SpMV (pnrm, nrow of them): every block sums them in the same order and
// takes the same decision; block 0 records it in the control block.
template <typename TV, int PASS, int CGS_PER>
__global__ void __launch_bounds__(B) k_cgs(int n, int k, TV *__restrict__ vk,
                                          const TV *__restrict__ V, KCtl *C, double *part,
                                          const double *pnrm, int nrow)
{
  if(gated(C, 2) || C->ierr12) return;
  if(PASS == 1) {
    __shared__ int s_bad;
    const double tot = sum_partials<double, B>(pnrm, nrow);
    if(threadIdx.x == 0) {
      s_bad = !(tot > 0.0 && 1.0 / tot > 0.0);
      if(blockIdx.x == 0) fin_nrm0(C, tot);
    }
    __syncthreads();
    if(s_bad) return;
  }
  const int nb = gridDim.x;
  const long base = (long)blockIdx.x * B * CGS_PER;
  double x[CGS_PER];
  long jj[CGS_PER];
#pragma unroll
  for(int u = 0; u < CGS_PER; u++) {
    jj[u] = base + u * B + threadIdx.x;
    x[u] = jj[u] < n ? (double)vk[jj[u]] : 0.;
  }
  if(PASS >= 2) {
    const double *hsub = (PASS == 2) ? C->h1 : C->h2;
    for(int i = 0; i < k; i++) {
      const double h = hsub[i];
      const TV *vi = V + (long)i * n;
#pragma unroll
      for(int u = 0; u < CGS_PER; u++)
        if(jj[u] < n) x[u] -= h * (double)vi[jj[u]];
    }
#pragma unroll
    for(int u = 0; u < CGS_PER; u++)
      if(jj[u] < n) {
        vk[jj[u]] = (TV)x[u];
        x[u] = (double)(TV)x[u];
      }
  }
  if(PASS <= 2) {
    // the k block sums without a barrier per coefficient: each wave reduces
    // its lanes (the tree of block_sum), the waves' sums meet in LDS and are
    // added in wave order -- the same arithmetic as block_sum
    __shared__ double red[kMaxKrylov + 2][B / kWave];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for(int i = 0; i < k; i++) {
      const TV *vi = V + (long)i * n;
      double a = 0.;
#pragma unroll
      for(int u = 0; u < CGS_PER; u++)
        if(jj[u] < n) a += (double)vi[jj[u]] * x[u];
      a = wave_sum(a);
      if(lane == 0) red[i][wv] = a;
    }
    __syncthreads();
    for(int i = threadIdx.x; i < k; i += B) {
      double t = 0.;
#pragma unroll
      for(int w = 0; w < B / kWave; w++) t += red[i][w];
      part[(long)i * nb + blockIdx.x] = t;
    }
  }
  else {
    double ss = 0.;
#pragma unroll
    for(int u = 0; u < CGS_PER; u++) ss += x[u] * x[u];
    double t = block_sum<double, B>(ss);
    if(threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// CGS2 finalisation, one block of 1024 threads.  Passes 1, 2: h (h') into the
// control block, one wave per coefficient.  Pass 3: the Hessenberg column and
// nrm1, then the end of the Arnoldi step (Givens rotation), which also runs
// when an abnormal |w| stopped the passes.
constexpr int FB = 1024;
template <int PASS>
__global__ void __launch_bounds__(FB) k_fin_cgs(KCtl *C, const double *part, int nparts, int k)
{
  if(gated(C, 2)) return;
  const bool stopped = C->ierr12 != 0;
  if(PASS <= 2) {
    if(!stopped) wave_sums<FB>(part, nparts, k, PASS == 1 ? C->h1 : C->h2);
    return;
  }
  double tot = 0.;
  if(!stopped) tot = sum_partials<double, FB>(part, nparts);
  if(threadIdx.x == 0) {
    if(!stopped) fin_hcol(C, k, tot, true);
    gm_givens_step(k, C);
  }
}

// CG: p = r = rhs - Ap (A*sol in ap); partial |r|^2
template <typename TV>
__global__ void __launch_bounds__(B) k_cg_init(int n, const TV *rhs, const TV *ap, TV *r, TV *p,
                                              double *part)
{
  double s = 0;
  for(long i = (long)blockIdx.x * B + threadIdx.x; i < n; i += (long)gridDim.x * B) {
    TV v = rhs[i] - ap[i];
    p[i] = v;
    r[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// CG: dx += alpha p, r -= alpha Ap; partial |r|^2
template <typename TV>
__global__ void __launch_bounds__(B) k_cg_update(int n, const TV *p, const TV *ap, TV *r, TV *dx,
                                                const KCtl *C, double *part)
{
  if(C->done) return;
  const double alpha = C->alpha;
  double s = 0;
  for(long i = (long)blockIdx.x * B + threadIdx.x; i < n; i += (long)gridDim.x * B) {
    dx[i] = (TV)((double)dx[i] + alpha * (double)p[i]);
    TV v = (TV)((double)r[i] - alpha * (double)ap[i]);
    r[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename TV>
__global__ void __launch_bounds__(B) k_cg_p(int n, const TV *r, TV *p, const KCtl *C)
{
  if(C->done) return;
  const double beta = C->beta;
  for(long i = (long)blockIdx.x * B + threadIdx.x; i < n; i += (long)gridDim.x * B)
    p[i] = (TV)((double)r[i] + beta * (double)p[i]);
}

// ---- CG without separate finalize kernels -------------------------------------------------
// Each block of the next streaming kernel sums the previous kernel's partials
// itself (same fixed order in every block, hence the same value), takes the
// scalar decision, and block 0 alone records it.  <r,r> of the previous
// iteration is read while the new one is written, so it lives in two slots
// selected by the iteration's parity (rr[2]).


// dx += alpha p, r -= alpha Ap with alpha = rr_old / <p, Ap> (iters.f:478-486);
// partial |r|^2.
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_cg_update2(int n, const TV *p, const TV *ap, TV *r, TV *dx,
                                                  KCtl *C, const double *part_pap, int npap,
                                                  double *part, int par)
{
  __shared__ double s_alpha;
  __shared__ int s_skip;
  if(C->done) return;
  double pap = sum_partials<double, BS>(part_pap, npap);
  if(threadIdx.x == 0) {
    int bk = 0;  // brkdn, iters.f:3368-3401
    if(pap > 0.0) bk = (1.0 / pap > 0.0) ? 0 : -9;
    else if(pap < 0.0) bk = (1.0 / pap < 0.0) ? 0 : -9;
    else if(pap == 0.0) bk = -3;
    else bk = -9;
    s_skip = bk;
    s_alpha = bk ? 0. : C->rr[par] / pap;
    if(blockIdx.x == 0) {
      C->its++;
      if(bk) {
        C->status = bk;
        C->done = 1;
      }
      else
        C->alpha = s_alpha;
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double alpha = s_alpha;
  double s = 0;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
    dx[i] = (TV)((double)dx[i] + alpha * (double)p[i]);
    TV v = (TV)((double)r[i] - alpha * (double)ap[i]);
    r[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<double, BS>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// stopbis on |r| (iters.f:3275-3330), then p = r + beta p.
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_cg_p2(int n, const TV *r, TV *p, KCtl *C,
                                             const double *part_rr, int nrr, int par)
{
  __shared__ double s_beta;
  __shared__ int s_stop;
  if(C->done) return;
  double rr = sum_partials<double, BS>(part_rr, nrr);
  if(threadIdx.x == 0) {
    int stop = 0, status = 0;
    double fpar5 = 0., beta = 0.;
    if(C->maxits > 0 && C->its >= C->maxits) {
      stop = 1;
      status = -1;
    }
    else {
      fpar5 = sqrt(rr);
      if(fpar5 > C->fpar4) beta = fpar5 * fpar5 / C->rr[par];
      else stop = 2;
    }
    s_stop = stop;
    s_beta = beta;
    if(blockIdx.x == 0) {
      if(stop == 1) {
        C->status = status;
        C->done = 1;
      }
      else {
        C->fpar5 = fpar5;
        C->fpar6 = fpar5;
        C->rr[1 - par] = fpar5 * fpar5;
        C->fpar7 = fpar5 * fpar5;
        C->beta = beta;
        if(stop == 2) {
          C->ipar11 = 1;
          C->done = 1;
        }
      }
    }
  }
  __syncthreads();
  if(s_stop) return;
  const double beta = s_beta;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS)
    p[i] = (TV)((double)r[i] + beta * (double)p[i]);
}

// ---- BCGSTAB (iters.f:1032-1306) ----------------------------------------------------------
// Same scheme as the CG kernels: every block re-derives the scalar of the
// step from the previous kernel's partials, block 0 records it.  Vectors in
// the reference's work columns: r0 = w1, r = w2, s = w3, t = w4, v = w5,
// p = w6, dx = w8 (w7, the preconditioner's output, is a separate z).

// brkdn (iters.f:3368-3401): 0, -3 (zero) or -9 (abnormal number)
__device__ inline int brkdn_code(double a)
{
  if(a > 0.0) return (1.0 / a > 0.0) ? 0 : -9;
  if(a < 0.0) return (1.0 / a < 0.0) ? 0 : -9;
  if(a == 0.0) return -3;
  return -9;
}

// labels 10/20: r0 = r = p = rhs - A sol (A sol in t); partial |r0|^2
template <typename TV>
__global__ void __launch_bounds__(B) k_bs_init(int n, const TV *rhs, const TV *t, TV *r0, TV *r, TV *p,
                                              double *part)
{
  double s = 0;
  for(long i = (long)blockIdx.x * B + threadIdx.x; i < n; i += (long)gridDim.x * B) {
    TV v = rhs[i] - t[i];
    r0[i] = v;
    r[i] = v;
    p[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<double, B>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// t = A x fused with the partial sums of <t,t> (part) and <t,s> (part + np).
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) k_op_dot2(SellOp<TA> A, const TV *__restrict__ x,
                                                 const TV *__restrict__ sv, TV *__restrict__ t,
                                                 const KCtl *C, double *part, int np)
{
  if(C->done) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  double tt = 0., ts = 0.;
  if(row < A.n) {
    double sum;
    sell_row<TA, TV>(row, A, x, sum);
    TV tv = (TV)sum;
    t[row] = tv;
    tt = (double)tv * (double)tv;
    ts = (double)tv * (double)sv[row];
  }
  tt = block_sum<double, 256>(tt);
  ts = block_sum<double, 256>(ts);
  if(threadIdx.x == 0) {
    part[blockIdx.x] = tt;
    part[np + blockIdx.x] = ts;
  }
}

// steps (2), (3): alpha = rho / <r0, v> (label 60), s = r - alpha v
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_bs_s(int n, const TV *r, const TV *v, TV *s, KCtl *C,
                                            const double *part_rv, int np, int par)
{
  __shared__ double s_alpha;
  __shared__ int s_skip;
  if(C->done) return;
  double den = sum_partials<double, BS>(part_rv, np);
  if(threadIdx.x == 0) {
    int bk = brkdn_code(den);
    s_skip = bk;
    s_alpha = bk ? 0. : C->rr[par] / den;
    if(blockIdx.x == 0) {
      C->its++;
      if(bk) {
        C->status = bk;
        C->done = 1;
      }
      else
        C->alpha = s_alpha;
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double alpha = s_alpha;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS)
    s[i] = (TV)((double)r[i] - alpha * (double)v[i]);
}

// step (5): omega = <t,s> / <t,t> (label 90) with both breakdown tests;
// steps (6), (7): dx += alpha p + omega s, r = s - omega t; partial |r|^2
// (part) and <r, r0> (part + np) for the stopping test and step (8).
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_bs_update(int n, const TV *p, const TV *s, const TV *t,
                                                 const TV *r0, TV *r, TV *dx, KCtl *C,
                                                 const double *part_t, int npt, double *part, int np)
{
  __shared__ double s_omega;
  __shared__ int s_skip;
  if(C->done) return;
  double tt = sum_partials<double, BS>(part_t, npt);
  double ts = sum_partials<double, BS>(part_t + npt, npt);
  if(threadIdx.x == 0) {
    int bk = brkdn_code(tt);
    double omega = 0.;
    if(!bk) {
      omega = ts / tt;
      bk = brkdn_code(omega);
    }
    s_skip = bk;
    s_omega = omega;
    if(blockIdx.x == 0) {
      C->its++;
      if(bk) {
        C->status = bk;
        C->done = 1;
      }
      else
        C->omega = omega;
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double alpha = C->alpha, omega = s_omega;
  double rr = 0., rr0 = 0.;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
    const double si = (double)s[i];
    const double w7 = alpha * (double)p[i] + omega * si;
    dx[i] = (TV)((double)dx[i] + w7);
    TV v = (TV)(si - omega * (double)t[i]);
    r[i] = v;
    rr += (double)v * (double)v;
    rr0 += (double)v * (double)r0[i];
  }
  rr = block_sum<double, BS>(rr);
  rr0 = block_sum<double, BS>(rr0);
  if(threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[np + blockIdx.x] = rr0;
  }
}

// stopbis on |r| (iters.f:3275-3330), then step (8): rho = <r, r0>,
// beta = rho alpha / (omega rho_old), p = r + beta (p - omega v), brkdn(rho).
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_bs_p(int n, const TV *r, const TV *v, TV *p, KCtl *C,
                                            const double *part, int np, int par)
{
  __shared__ double s_beta;
  __shared__ int s_stop;
  if(C->done) return;
  double rr = sum_partials<double, BS>(part, np);
  double rho = sum_partials<double, BS>(part + np, np);
  if(threadIdx.x == 0) {
    int stop = 0;
    double fpar5 = 0., beta = 0.;
    if(C->maxits > 0 && C->its >= C->maxits) stop = 1;
    else {
      fpar5 = sqrt(rr);
      if(fpar5 > C->fpar4) beta = rho * C->alpha / (C->omega * C->rr[par]);
      else stop = 2;
    }
    s_stop = stop;
    s_beta = beta;
    if(blockIdx.x == 0) {
      if(stop == 1) {
        C->status = -1;
        C->done = 1;
      }
      else {
        C->fpar5 = fpar5;
        C->fpar6 = fpar5;
        if(stop == 2) {
          C->ipar11 = 1;
          C->done = 1;
        }
        else {
          C->rr[1 - par] = rho;
          C->fpar7 = rho;
          C->beta = beta;
          const int bk = brkdn_code(rho);
          if(bk) {  // tested after p is formed; the solve ends here
            C->status = bk;
            C->done = 1;
          }
        }
      }
    }
  }
  __syncthreads();
  if(s_stop) return;
  const double beta = s_beta, omega = C->omega;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS)
    p[i] = (TV)((double)r[i] + beta * ((double)p[i] - omega * (double)v[i]));
}

// ---- TFQMR (iters.f:1307-1584) ------------------------------------------------------------
// Work columns of the reference: r0 = w2, w = w3, y = w4, z = w5, ay = w6,
// az = w7, v = w8, d = w9, dx = w11 (w10, the preconditioner's output, is z
// of the driver).  Scalars: rho in parity slots rr[], beta (the routine's
// first `alpha`), alpha, tau and te after update I (tao1, te1) and II (tao, te).
// Update II both updates dx and may end the solve, so its block 0 only
// raises pend_flag; the next kernel turns it into `done` (all its blocks
// skip), and no block of update II can miss its share of the update.

// labels 10/20: r0 = w = rhs - A sol (A sol in az); partial <r0, w> for
// rho of the first iteration; the scalars bisinit clears.
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_tq_init(int n, const TV *rhs, const TV *az, TV *r0, TV *w, KCtl *C,
                                               double *part)
{
  double s = 0;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
    TV v = rhs[i] - az[i];
    r0[i] = v;
    w[i] = v;
    s += (double)v * (double)v;
  }
  s = block_sum<double, BS>(s);
  if(threadIdx.x == 0) {
    part[blockIdx.x] = s;
    if(blockIdx.x == 0) {
      C->its = 1;
      C->status = 1;
      C->te = 0.;
      C->rr[0] = 0.;
      C->fpar6 = 0.;
      C->pend_flag = 0;
    }
  }
}

// label 30: rho = <r0, w>, brkdn(rho), beta = rho / rho_old (0 on the first
// iteration), y = w + beta z.  The first iteration also sets tau = |r0| and
// the target fpar(4) (label 20).
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_tq_y(int n, const TV *w, const TV *z, TV *y, KCtl *C,
                                            const double *part_rho, int np, int par)
{
  __shared__ double s_beta;
  __shared__ int s_skip;
  if(C->done) return;
  if(C->pend_flag) {
    if(blockIdx.x == 0 && threadIdx.x == 0) C->done = 1;
    return;
  }
  double rho = sum_partials<double, BS>(part_rho, np);
  if(threadIdx.x == 0) {
    const int bk = brkdn_code(rho);
    const bool first = C->its == 1;
    s_skip = bk;
    s_beta = (bk || first) ? 0. : rho / C->rr[par];
    if(blockIdx.x == 0) {
      if(first) {
        C->fpar5 = sqrt(rho);
        C->fpar3 = C->fpar5;
        C->tao = C->fpar5;
        C->fpar4 = C->rtol * C->tao + 0.0;
      }
      if(bk) {
        C->status = bk;
        C->done = 1;
      }
      else {
        C->rr[1 - par] = rho;
        C->beta = s_beta;
      }
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double beta = s_beta;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS)
    y[i] = (TV)((double)w[i] + beta * (double)z[i]);
}

// ay = A x, v = ay + beta (az + beta v); partial <r0, v>.
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) k_tq_op_v(SellOp<TA> A, const TV *__restrict__ x,
                                                 const TV *__restrict__ az, const TV *__restrict__ r0,
                                                 TV *__restrict__ ay, TV *__restrict__ v, const KCtl *C,
                                                 double *part)
{
  if(C->done) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  double t = 0.;
  if(row < A.n) {
    double sum;
    sell_row<TA, TV>(row, A, x, sum);
    TV ayv = (TV)sum;
    ay[row] = ayv;
    const double beta = C->beta;
    TV vv = (TV)((double)ayv + beta * ((double)az[row] + beta * (double)v[row]));
    v[row] = vv;
    t = (double)r0[row] * (double)vv;
  }
  t = block_sum<double, 256>(t);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

// label 60: sigma = <r0, v>, brkdn, alpha = rho / sigma, z = y - alpha v.
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_tq_z(int n, const TV *y, const TV *v, TV *z, KCtl *C,
                                            const double *part_sig, int np, int par)
{
  __shared__ double s_alpha;
  __shared__ int s_skip;
  if(C->done) return;
  double sigma = sum_partials<double, BS>(part_sig, np);
  if(threadIdx.x == 0) {
    const int bk = brkdn_code(sigma);
    s_skip = bk;
    s_alpha = bk ? 0. : C->rr[1 - par] / sigma;
    if(blockIdx.x == 0) {
      C->its++;
      if(bk) {
        C->status = bk;
        C->done = 1;
      }
      else
        C->alpha = s_alpha;
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double alpha = s_alpha;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS)
    z[i] = (TV)((double)y[i] - alpha * (double)v[i]);
}

// az = A x, w = w - alpha ay (label 90); partial |w|^2.
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) k_tq_op_w(SellOp<TA> A, const TV *__restrict__ x,
                                                 const TV *__restrict__ ay, TV *__restrict__ az,
                                                 TV *__restrict__ w, const KCtl *C, double *part)
{
  if(C->done) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  double t = 0.;
  if(row < A.n) {
    double sum;
    sell_row<TA, TV>(row, A, x, sum);
    az[row] = (TV)sum;
    TV wv = (TV)((double)w[row] - C->alpha * (double)ay[row]);
    w[row] = wv;
    t = (double)wv * (double)wv;
  }
  t = block_sum<double, 256>(t);
  if(threadIdx.x == 0) part[blockIdx.x] = t;
}

// tau update shared by updates I and II: theta = |w|^2 / tau^2,
// sigma = 1 / (1 + theta), tau' = tau sqrt(sigma theta), eta = sigma alpha,
// the coefficient of d te / alpha, te' = theta eta.
__device__ inline void tq_tau(double ww, double tao, double te, double alpha, double &tao_new,
                              double &eta, double &dcoef, double &te_new)
{
  const double theta = ww / (tao * tao);
  const double sigma = 1.0 / (1.0 + theta);
  tao_new = tao * sqrt(sigma * theta);
  eta = sigma * alpha;
  dcoef = te / alpha;
  te_new = theta * eta;
}

// update I: d = y + (te/alpha) d, dx += eta d, w = w - alpha az; partial |w|^2
// (part) and <r0, w> (part + np, rho of the next iteration).
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_tq_upd1(int n, const TV *y, const TV *az, const TV *r0, TV *d,
                                               TV *dx, TV *w, KCtl *C, const double *part_w, int npw,
                                               double *part, int np)
{
  __shared__ double s_eta, s_dc;
  __shared__ int s_skip;
  if(C->done) return;
  double ww = sum_partials<double, BS>(part_w, npw);
  if(threadIdx.x == 0) {
    double tao, eta, dc, te;
    tq_tau(ww, C->tao, C->te, C->alpha, tao, eta, dc, te);
    const int bk = brkdn_code(tao);
    s_skip = bk;
    s_eta = eta;
    s_dc = dc;
    if(blockIdx.x == 0) {
      C->its++;
      if(bk) {
        C->status = bk;
        C->done = 1;
      }
      else {
        C->tao1 = tao;
        C->te1 = te;
      }
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double eta = s_eta, dc = s_dc, alpha = C->alpha;
  double s1 = 0., s2 = 0.;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
    const double di = (double)y[i] + dc * (double)d[i];
    const TV dv = (TV)di;
    d[i] = dv;
    dx[i] = (TV)((double)dx[i] + eta * (double)dv);
    const TV wv = (TV)((double)w[i] - alpha * (double)az[i]);
    w[i] = wv;
    s1 += (double)wv * (double)wv;
    s2 += (double)r0[i] * (double)wv;
  }
  s1 = block_sum<double, BS>(s1);
  s2 = block_sum<double, BS>(s2);
  if(threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[np + blockIdx.x] = s2;
  }
}

// update II: d = z + (te/alpha) d, dx += eta d; fpar(5) = fpar(6) = tau and
// the loop test of label 90 (tolerance or iteration limit).
template <typename TV, int BS>
__global__ void __launch_bounds__(BS) k_tq_upd2(int n, const TV *z, TV *d, TV *dx, KCtl *C,
                                               const double *part_w, int npw)
{
  __shared__ double s_eta, s_dc;
  __shared__ int s_skip;
  if(C->done) return;
  double ww = sum_partials<double, BS>(part_w, npw);
  if(threadIdx.x == 0) {
    double tao, eta, dc, te;
    tq_tau(ww, C->tao1, C->te1, C->alpha, tao, eta, dc, te);
    const int bk = brkdn_code(tao);
    s_skip = bk;
    s_eta = eta;
    s_dc = dc;
    if(blockIdx.x == 0) {
      if(bk) {
        C->status = bk;
        C->pend_flag = 1;
      }
      else {
        C->tao = tao;
        C->te = te;
        C->fpar5 = tao;
        C->fpar6 = tao;
        if(!(tao > C->fpar4 && (C->its < C->maxits || C->maxits <= 0))) C->pend_flag = 1;
      }
    }
  }
  __syncthreads();
  if(s_skip) return;
  const double eta = s_eta, dc = s_dc;
  for(long i = (long)blockIdx.x * BS + threadIdx.x; i < n; i += (long)gridDim.x * BS) {
    const TV dv = (TV)((double)z[i] + dc * (double)d[i]);
    d[i] = dv;
    dx[i] = (TV)((double)dx[i] + eta * (double)dv);
  }
}

// ---- finalize kernels (one block) -------------------------------------------------------

enum FinOp {
  FIN_GM_RESID,   // gmres label 10/20: alpha = |w|, first tolerance, early exit
  FIN_NRM0,       // mgsro: nrm0 = |v|^2, abnormal-number checks
  FIN_MGS_DOT,    // h_i = V_k . V_i
  FIN_MGS_AXPY,   // re-orthogonalisation test, running norm update
  FIN_MGS_NORM,   // nrm1 = |V_k|, Hessenberg column
  FIN_CG_INIT,
  FIN_CG_PAP,     // brkdn + alpha = <r,r>/<p,Ap>
  FIN_CG_UPDATE,  // stopbis (iters.f:3275-3330)
  FIN_BS_INIT     // bcgstab label 20: rho, tolerance, early exit (iters.f:1140-1154)
};

template <int OP>
__global__ void __launch_bounds__(B) k_fin(KCtl *C, const double *part, int nparts, int idx)
{
  // same gate as the producing kernel
  if(OP == FIN_GM_RESID || OP == FIN_CG_PAP || OP == FIN_CG_UPDATE) {
    if(C->done) return;
  }
  else if(OP == FIN_NRM0) {
    if(gated(C, 2)) return;
  }
  else if(OP != FIN_CG_INIT && OP != FIN_BS_INIT) {
    if(gated(C, 2) || C->ierr12) return;
  }
  const double tot = sum_partials<double, B>(part, nparts);
  if(threadIdx.x) return;
  if(OP == FIN_GM_RESID) {
    double alpha = sqrt(tot);
    C->its++;
    if(C->its == 1) {
      C->fpar4 = C->rtol * alpha + 0.0;
      C->fpar3 = alpha;
    }
    C->fpar5 = alpha;
    C->g[0] = alpha;
    C->cycle_end = 0;
    C->kfinal = 0;
    C->ierr12 = 0;
    C->pend_flag = 0;
    if(alpha <= C->fpar4) {
      C->status = 0;
      C->fpar6 = alpha;
      C->done = 1;
    }
    else
      C->scal = 1.0 / alpha;
  }
  else if(OP == FIN_NRM0) {
    fin_nrm0(C, tot);
  }
  else if(OP == FIN_MGS_DOT) {
    C->h1[idx] = tot;
    C->pend_flag = 0;
  }
  else if(OP == FIN_MGS_AXPY) {
    const double f = C->h1[idx];
    if(f * f > C->thr) {
      C->h1[idx] = C->h1[idx] + tot;
      C->pend = tot;
      C->pend_flag = 1;
    }
    else
      C->pend_flag = 0;
    double nrm0 = C->nrm0 - C->h1[idx] * C->h1[idx];
    if(nrm0 < 0.0) nrm0 = 0.0;
    C->nrm0 = nrm0;
    C->thr = nrm0 * 0.98;
  }
  else if(OP == FIN_MGS_NORM) {
    fin_hcol(C, idx, tot, false);
  }
  else if(OP == FIN_CG_INIT) {
    C->its = 1;
    C->fpar7 = tot;
    C->rr[0] = tot;
    C->fpar3 = sqrt(tot);
    C->fpar5 = C->fpar3;
    C->fpar4 = C->rtol * C->fpar3 + 0.0;
    C->fpar6 = 0.0;
    C->ipar11 = 0;
    C->status = 1;
  }
  else if(OP == FIN_BS_INIT) {
    C->its = 1;
    C->fpar7 = tot;
    C->rr[0] = tot;
    C->fpar5 = sqrt(tot);
    C->fpar3 = C->fpar5;
    C->fpar4 = C->rtol * C->fpar3 + 0.0;
    C->fpar6 = C->fpar5;
    C->ipar11 = 0;
    C->status = 1;
    C->done = C->fpar5 <= C->fpar4 ? 1 : 0;
  }
  else if(OP == FIN_CG_PAP) {
    const double alpha = tot;
    C->its++;
    int bk = 0;  // brkdn, iters.f:3368-3401
    if(alpha > 0.0) bk = (1.0 / alpha > 0.0) ? 0 : -9;
    else if(alpha < 0.0) bk = (1.0 / alpha < 0.0) ? 0 : -9;
    else if(alpha == 0.0) bk = -3;
    else bk = -9;
    if(bk) {
      C->status = bk;
      C->done = 1;
    }
    else
      C->alpha = C->fpar7 / alpha;
  }
  else if(OP == FIN_CG_UPDATE) {
    if(C->maxits > 0 && C->its >= C->maxits) {
      C->status = -1;
      C->done = 1;
      return;
    }
    C->fpar5 = sqrt(tot);
    C->fpar6 = C->fpar5;
    if(C->fpar6 > C->fpar4) {
      C->ipar11 = 0;
      C->beta = C->fpar5 * C->fpar5 / C->fpar7;
      C->fpar7 = C->fpar5 * C->fpar5;
    }
    else {
      C->ipar11 = 1;
      C->done = 1;
    }
  }
}

__global__ void k_gm_givens(int k, KCtl *C)
{
  if(threadIdx.x || blockIdx.x || gated(C, 2)) return;
  gm_givens_step(k, C);
}

// label 200: back substitution of the rotated Hessenberg system.
__global__ void k_gm_backsub(KCtl *C)
{
  if(threadIdx.x || blockIdx.x || C->done) return;
  int k = C->kfinal;
  while(k > 0 && C->H[(long)(k - 1) * LDH + (k - 1)] == 0.0) k--;
  if(k <= 0) {
    C->status = -3;
    C->ierr12 = -4;
    C->done = 1;
    C->kfinal = 0;
    return;
  }
  C->g[k - 1] = C->g[k - 1] / C->H[(long)(k - 1) * LDH + (k - 1)];
  for(int i = k - 1; i >= 1; i--) {
    const double *hcol = C->H + (long)i * LDH;
    for(int ii = 0; ii < i; ii++) C->g[ii] = C->g[ii] - C->g[i] * hcol[ii];
    C->g[i - 1] = C->g[i - 1] / C->H[(long)(i - 1) * LDH + (i - 1)];
  }
  C->kfinal = k;
}

// GMRES update vector u = sum_i y_i V_i, accumulated into V_0 as the reference
// does (w(ii) = w(ii)*y_1, then w += y_i v_i).
template <typename TV>
__global__ void __launch_bounds__(B) k_gm_combine(int n, TV *V, const KCtl *C)
{
  if(C->done) return;
  const int k = C->kfinal;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) {
    double u = (double)V[j] * C->g[0];
    for(int i = 1; i < k; i++) u = u + C->g[i] * (double)V[(long)i * n + j];
    V[j] = (TV)u;
  }
}

// FGMRES update: sol += y_i Z_i, basis by basis.
template <typename TV>
__global__ void __launch_bounds__(B) k_fgm_update(int n, TV *sol, const TV *Z, const KCtl *C)
{
  if(C->done) return;
  const int k = C->kfinal;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) {
    double x = (double)sol[j];
    for(int i = 0; i < k; i++) x = x + C->g[i] * (double)Z[(long)i * n + j];
    sol[j] = (TV)x;
  }
}

template <typename TV>
__global__ void __launch_bounds__(B) k_axpy1(int n, TV *sol, const TV *d, const KCtl *C)
{
  if(C->done) return;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B)
    sol[j] = sol[j] + d[j];
}

template <typename TV>
__global__ void __launch_bounds__(B) k_copy_gated(int n, const TV *src, TV *dst, const KCtl *C,
                                                 int gmode)
{
  if(gated(C, gmode)) return;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) dst[j] = src[j];
}

// label 60/70: restart or terminate.
__global__ void k_gm_cycle_end(KCtl *C)
{
  if(threadIdx.x || blockIdx.x || C->done) return;
  C->fpar6 = C->fpar5;
  if(C->ierr12 != 0) {
    C->status = -3;
    C->done = 1;
    return;
  }
  if((C->its < C->maxits || C->maxits <= 0) && C->fpar6 > C->fpar4) return;  // restart
  if(C->fpar6 <= C->fpar4) C->status = 0;
  else if(C->its >= C->maxits && C->maxits > 0) C->status = -1;
  else C->status = -10;
  C->done = 1;
}

int reduce_blocks(int n)
{
  long nb = ((long)n + B * 4 - 1) / (B * 4);
  return (int)std::max(1L, std::min(nb, 1024L));
}

struct Workspace {
  KCtl *ctl = nullptr;
  double *part = nullptr;
  void *vec = nullptr;
  size_t vec_bytes = 0, part_size = 0;
  ~Workspace()
  {
    dfree(ctl);
    dfree(part);
    if(vec) (void)hipFree(vec);
  }
};

Workspace &ws()
{
  static Workspace w;
  return w;
}

// rows of one partial-sum slot of the workspace
size_t part_rows(int n) { return (size_t)std::max(1024L, ((long)n + 255) / 256); }

int ensure_ws(size_t vec_bytes, int n)
{
  Workspace &W = ws();
  size_t rows = part_rows(n);
  size_t need = rows * (kMaxKrylov + 2);
  if(!W.ctl) GDP_TRY(dalloc(&W.ctl, 1));
  if(W.part_size < need) {
    dfree(W.part);
    GDP_TRY(dalloc(&W.part, need));
    W.part_size = need;
  }
  if(W.vec_bytes < vec_bytes) {
    if(W.vec) (void)hipFree(W.vec);
    W.vec = nullptr;
    W.vec_bytes = 0;
    GDP_HIP(hipMalloc(&W.vec, vec_bytes));
    W.vec_bytes = vec_bytes;
  }
  return GDP_OK;
}

template <typename TV>
int prec_apply(const PrecOp<TV> &M, const TV *in, TV *out, KCtl *C, int gmode, int n)
{
  if(M.kind == PREC_ILU) return M.ilu->template apply<TV>(in, (TV)1, out, (const int *)C, gmode);
  if(M.kind == PREC_USER) return user_prec_apply(M.user, in, out, C, gmode);
  if(M.kind == PREC_INNER32) {
    if constexpr(sizeof(TV) == 8) return inner_apply(M.inner, in, out, C, gmode);
    else return set_error(GDP_E_UNSUPPORTED, "nested FP32 preconditioner inside an FP32 solve");
  }
  dim3 g(reduce_blocks(n));
  hipLaunchKernelGGL(k_copy_gated<TV>, g, dim3(B), 0, ctx().stream, n, in, out, C, gmode);
  return GDP_OK;
}

// Two-stage flexible preconditioner: the learned map, then an ILU(0)
// correction of the residual it leaves, z = a N(v) + (LU)^{-1} (v - a A N(v)),
// with a = 1, or (network mode bit 1) the least-squares scale
// a = <v, A N(v)> / |A N(v)|^2 for networks whose output scale is not fixed
// by their training (trained on the direction of the error only).
template <typename TA, typename TV>
int prec_apply_two_stage(const DevOp<TA> &A, const PrecOp<TV> &M, const TV *in, TV *out, TV *t,
                         TV *u, double *part, KCtl *C, int gmode, int n)
{
  GDP_TRY(user_prec_apply(M.user, in, out, C, gmode));
  if(n && !user_prec_lsq(M.user))
    hipLaunchKernelGGL((k_op_resid<TA, TV>), dim3((n + 255) / 256), dim3(256), 0, ctx().stream, A, out, in, t, C,
                       gmode);
  else if(n) {
    // partial sums and the scale in the cycle's partial-sum buffer (at least
    // 66 ceil(n / 256) doubles), free while the step preconditions
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL((k_op_dots<TA, TV>), dim3(nb), dim3(256), 0, ctx().stream, A, out, in, t, part, C, gmode);
    hipLaunchKernelGGL(k_alpha, dim3(1), dim3(256), 0, ctx().stream, nb, part, C, gmode);
    hipLaunchKernelGGL((k_scale_resid<TV>), dim3(nb), dim3(256), 0, ctx().stream, n, part + 2 * nb, out, in, t, C,
                       gmode);
  }
  // u = (LU)^{-1} t, and out += u in the backward sweep
  GDP_TRY(M.ilu->template apply<TV>(t, (TV)1, u, (const int *)C, gmode, out));
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

template <int OP> void fin(KCtl *C, const double *part, int nparts, int idx)
{
  hipLaunchKernelGGL(k_fin<OP>, dim3(1), dim3(B), 0, ctx().stream, C, part, nparts, idx);
}

// One restart cycle of Sparskit's gmres/fgmres (iters.f:1912-2217,
// 2522-2807) with every decision taken on the device (control block C):
// residual, m Arnoldi steps (preconditioning, SpMV, Gram-Schmidt, Givens),
// back-substitution, update, restart/stop test.  Enqueue only.
template <typename TA, typename TV> struct GmArgs {
  DevOp<TA> A;
  const PrecOp<TV> *Mp;
  const TV *rhs;
  TV *sol, *V, *Z;
  KCtl *C;
  double *part;
  int n, m, orth;
  bool flexible, two_stage;
};

template <typename TA, typename TV> int gm_cycle(const GmArgs<TA, TV> &ga)
{
  const DevOp<TA> &A = ga.A;
  const PrecOp<TV> &M = *ga.Mp;
  const TV *rhs = ga.rhs;
  TV *sol = ga.sol, *V = ga.V, *Z = ga.Z, *tmpz = ga.Z;
  KCtl *C = ga.C;
  double *part = ga.part;
  const int n = ga.n, m = ga.m;
  const bool flexible = ga.flexible, two_stage = ga.two_stage, rp = M.kind != PREC_NONE;
  hipStream_t s = ctx().stream;
  const int nb = reduce_blocks(n);
  const int nrow = std::max(1, (n + 255) / 256);
  const int cper = cgs_per(n);
  const int ncgs = (int)std::max(1L, ((long)n + B * cper - 1) / (B * cper));
  // partial sums: CGS coefficient i at [i * ncgs, (i + 1) * ncgs), inside the
  // first kMaxKrylov slots of part_rows(n); the SpMV's |w|^2 in the slot after
  // them (the workspace holds kMaxKrylov + 2 slots, ensure_ws)
  double *pnrm = part + (size_t)kMaxKrylov * part_rows(n);
  dim3 g(nb), blk(B), grow(nrow);
  GDP_TRY(op_apply(A, sol, V, C, 1));
  hipLaunchKernelGGL(k_gm_resid<TV>, g, blk, 0, s, n, rhs, V, C, part);
  fin<FIN_GM_RESID>(C, part, nb, 0);
  hipLaunchKernelGGL(k_scale<TV>, g, blk, 0, s, n, V, C, 1);
  for(int k = 1; k <= m; k++) {
    TV *vk1 = V + (size_t)(k - 1) * n, *vk = V + (size_t)k * n;
    const TV *src = vk1;
    if(two_stage) {
      TV *zk = Z + (size_t)(k - 1) * n;
      TV *t1 = Z + (size_t)m * n, *t2 = t1 + n;
      GDP_TRY((prec_apply_two_stage<TA, TV>(A, M, vk1, zk, t1, t2, ga.part, C, 2, n)));
      src = zk;
    }
    else if(flexible) {
      TV *zk = Z + (size_t)(k - 1) * n;
      GDP_TRY(prec_apply(M, vk1, zk, C, 2, n));
      src = zk;
    }
    else if(rp) {
      GDP_TRY(prec_apply(M, vk1, tmpz, C, 2, n));
      src = tmpz;
    }
    // w = A z with |w|^2 partials (into pnrm, apart from the CGS partials)
    hipLaunchKernelGGL((k_op_dot<TA, TV, false>), grow, dim3(256), 0, s, A, src, (const TV *)nullptr,
                       vk, C, 2, pnrm);
    if(ga.orth == 1) {
      // CGS2: three passes with one-block finalisations; the third ends the
      // step (Hessenberg column, Givens rotation)
      dim3 gc2(ncgs), gf(1), bf(FB);
      if(cper == 4) {
        hipLaunchKernelGGL((k_cgs<TV, 1, 4>), gc2, blk, 0, s, n, k, vk, V, C, part, pnrm, nrow);
        hipLaunchKernelGGL(k_fin_cgs<1>, gf, bf, 0, s, C, part, ncgs, k);
        hipLaunchKernelGGL((k_cgs<TV, 2, 4>), gc2, blk, 0, s, n, k, vk, V, C, part, pnrm, nrow);
        hipLaunchKernelGGL(k_fin_cgs<2>, gf, bf, 0, s, C, part, ncgs, k);
        hipLaunchKernelGGL((k_cgs<TV, 3, 4>), gc2, blk, 0, s, n, k, vk, V, C, part, pnrm, nrow);
      }
      else {
        hipLaunchKernelGGL((k_cgs<TV, 1, 1>), gc2, blk, 0, s, n, k, vk, V, C, part, pnrm, nrow);
        hipLaunchKernelGGL(k_fin_cgs<1>, gf, bf, 0, s, C, part, ncgs, k);
        hipLaunchKernelGGL((k_cgs<TV, 2, 1>), gc2, blk, 0, s, n, k, vk, V, C, part, pnrm, nrow);
        hipLaunchKernelGGL(k_fin_cgs<2>, gf, bf, 0, s, C, part, ncgs, k);
        hipLaunchKernelGGL((k_cgs<TV, 3, 1>), gc2, blk, 0, s, n, k, vk, V, C, part, pnrm, nrow);
      }
      hipLaunchKernelGGL(k_fin_cgs<3>, gf, bf, 0, s, C, part, ncgs, k);
      // V_k = w / nrm1; skipped when the step ended the cycle (V_k is then
      // not used: the update combines V_0..V_{k-1} or Z_0..Z_{k-1})
      hipLaunchKernelGGL(k_scale<TV>, g, blk, 0, s, n, vk, C, 2);
    }
    else {
      fin<FIN_NRM0>(C, pnrm, nrow, 0);
      for(int i = 0; i < k; i++) {
        hipLaunchKernelGGL(k_mgs_dot<TV>, g, blk, 0, s, n, i, vk, V, C, part);
        fin<FIN_MGS_DOT>(C, part, nb, i);
        hipLaunchKernelGGL(k_mgs_axpy<TV>, g, blk, 0, s, n, i, vk, V, C, part);
        fin<FIN_MGS_AXPY>(C, part, nb, i);
      }
      hipLaunchKernelGGL(k_mgs_norm<TV>, g, blk, 0, s, n, k, vk, V, C, part);
      fin<FIN_MGS_NORM>(C, part, nb, k);
      hipLaunchKernelGGL(k_scale<TV>, g, blk, 0, s, n, vk, C, 2);
      hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(1), 0, s, k, C);
    }
  }
  hipLaunchKernelGGL(k_gm_backsub, dim3(1), dim3(1), 0, s, C);
  if(flexible)
    hipLaunchKernelGGL(k_fgm_update<TV>, g, blk, 0, s, n, sol, Z, C);
  else {
    hipLaunchKernelGGL(k_gm_combine<TV>, g, blk, 0, s, n, V, C);
    if(rp) {
      GDP_TRY(prec_apply(M, V, tmpz, C, 1, n));
      hipLaunchKernelGGL(k_axpy1<TV>, g, blk, 0, s, n, sol, tmpz, C);
    }
    else
      hipLaunchKernelGGL(k_axpy1<TV>, g, blk, 0, s, n, sol, V, C);
  }
  hipLaunchKernelGGL(k_gm_cycle_end, dim3(1), dim3(1), 0, s, C);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

// Record or replay: enqueue fn() into a graph once, then replay it.
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  ~GraphCache()
  {
    if(exec) (void)hipGraphExecDestroy(exec);
    if(graph) (void)hipGraphDestroy(graph);
  }
};

// GDP_NO_GRAPHS=1 in the environment launches the cycle's kernels directly
// (for tools that trace individual dispatches).
bool graphs_disabled()
{
  static int v = -1;
  if(v < 0) {
    const char *e = std::getenv("GDP_NO_GRAPHS");
    v = (e && e[0] && e[0] != '0') ? 1 : 0;
  }
  return v == 1;
}

template <typename Fn> int run_maybe_graph(bool use, GraphCache &gc, Fn fn)
{
  hipStream_t s = ctx().stream;
  if(!use || graphs_disabled()) return fn();
  if(!gc.exec) {
    GDP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = fn();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if(r) {
      if(g) (void)hipGraphDestroy(g);
      return r;
    }
    if(e != hipSuccess) return set_error(GDP_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    gc.graph = g;
    GDP_HIP(hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0));
  }
  GDP_HIP(hipGraphLaunch(gc.exec, s));
  return GDP_OK;
}

int read_ctl(KCtl *d, int *done, int *status, int *its)
{
  int h[3];
  GDP_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx().stream));
  GDP_HIP(hipStreamSynchronize(ctx().stream));
  *done = h[0];
  *status = h[1];
  *its = h[2];
  return GDP_OK;
}

// tidycg (iters.f:3333-3365) after label 900 of cg / bcgstab / tfqmr: the final
// status, then sol += M^{-1} dx (right preconditioning) or sol += dx.
template <typename TV>
int tidy_cg(const PrecOp<TV> &M, KCtl *C, TV *dx, TV *z, TV *sol, int n, KrylovResult &res)
{
  hipStream_t s = ctx().stream;
  KCtl tmp;
  GDP_TRY(d2h(&tmp, C, 1));
  int final_status = tmp.status;
  if(final_status > 0) {
    if(tmp.fpar6 <= tmp.fpar4) final_status = 0;
    else if(tmp.its >= tmp.maxits && tmp.maxits > 0) final_status = -1;
    else final_status = -10;
  }
  int zero = 0;
  GDP_TRY(h2d(&C->done, &zero, 1));
  dim3 g(reduce_blocks(n)), blk(B);
  if(M.kind != PREC_NONE) {
    GDP_TRY(prec_apply(M, dx, z, C, 0, n));
    hipLaunchKernelGGL(k_axpy1<TV>, g, blk, 0, s, n, sol, z, C);
  }
  else
    hipLaunchKernelGGL(k_axpy1<TV>, g, blk, 0, s, n, sol, dx, C);
  GDP_HIP(hipGetLastError());
  GDP_TRY(sync());
  if(M.kind == PREC_ILU) GDP_TRY(M.ilu->check_timeout());
  res.its = tmp.its;
  res.status = final_status;
  res.res0 = tmp.fpar3;
  res.res_est = tmp.fpar6;
  return GDP_OK;
}

}  // namespace

template <typename TA, typename TV>
int krylov_solve(const DevOp<TA> &A, const PrecOp<TV> &M, const TV *rhs, TV *sol,
                 const KrylovOptions &opt, KrylovResult &res)
{
  const int n = A.n;
  hipStream_t s = ctx().stream;
  const int m = std::max(2, std::min(opt.m <= 1 ? 15 : opt.m, kMaxKrylov));
  const bool flexible = opt.algorithm == 9;
  const bool cg = opt.algorithm == 1;
  const bool bcg = opt.algorithm == 5;
  const bool tfq = opt.algorithm == 6;
  if(!cg && !bcg && !tfq && opt.algorithm != 8 && !flexible)
    return set_error(GDP_E_UNSUPPORTED, "Unknown algorithm for sparse matrix solver (%d)", opt.algorithm);
  const bool two_stage = flexible && M.kind == PREC_USER && M.ilu != nullptr;
  size_t nvec = cg ? 6 : bcg ? 8 : tfq ? 10 : (size_t)(m + 2) + (flexible ? m : 1) + (two_stage ? 1 : 0);
  GDP_TRY(ensure_ws(nvec * (size_t)std::max(1, n) * sizeof(TV) + 256, n));
  Workspace &W = ws();
  TV *base = (TV *)W.vec;
  KCtl hc;
  std::memset(&hc, 0, sizeof(hc));
  hc.rtol = opt.rtol;
  hc.maxits = opt.maxits;
  hc.m = m;
  GDP_TRY(h2d(W.ctl, &hc, 1));
  GDP_HIP(hipMemsetAsync(sol, 0, sizeof(TV) * n, s));  // x0 = 0 (Sparskit.cpp:226)
  KCtl *C = W.ctl;
  double *part = W.part;
  const int nb = reduce_blocks(n);
  const int nrow = std::max(1, (n + 255) / 256);
  const int cper = cgs_per(n);
  const int ncgs = (int)std::max(1L, ((long)n + B * cper - 1) / (B * cper));
  dim3 g(nb), blk(B), grow(nrow);
  int done = 0, status = 0, its = 0;

  if(cg) {
    TV *r = base, *p = base + n, *ap = base + 2 * (size_t)n, *z = base + 3 * (size_t)n,
       *dx = base + 4 * (size_t)n;
    GDP_HIP(hipMemsetAsync(dx, 0, sizeof(TV) * n, s));
    GDP_TRY(op_apply(A, sol, ap, C, 0));
    hipLaunchKernelGGL(k_cg_init<TV>, g, blk, 0, s, n, rhs, ap, r, p, part);
    fin<FIN_CG_INIT>(C, part, nb, 0);
    const bool rp = M.kind != PREC_NONE;
    GraphCache gc;
    const int batch = 16;
    // three kernels per iteration: SpMV + <p,Ap> partials, update + |r|^2
    // partials (each block derives alpha), p update (each block derives beta)
    // one workgroup of 1024 threads per CU for the two vector kernels: 16
    // waves keep enough loads in flight to stream the vectors
    constexpr int CGB = 1024;
    const int ncg = std::max(1, std::min(ctx().cu_count, (n + CGB - 1) / CGB));
    double *part2 = part + nrow;
    auto body = [&]() -> int {
      for(int it = 0; it < batch; it++) {
        const TV *src = p;
        if(rp) {
          GDP_TRY(prec_apply(M, p, z, C, 1, n));
          src = z;
        }
        hipLaunchKernelGGL((k_op_dot<TA, TV, true>), grow, dim3(256), 0, s, A, src, p, ap, C, 1, part);
        hipLaunchKernelGGL((k_cg_update2<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, p, ap, r, dx, C, part,
                           nrow, part2, it & 1);
        hipLaunchKernelGGL((k_cg_p2<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, r, p, C, part2, ncg, it & 1);
      }
      GDP_HIP(hipGetLastError());
      return GDP_OK;
    };
    for(;;) {
      GDP_TRY(run_maybe_graph(opt.use_graphs, gc, body));
      GDP_TRY(read_ctl(C, &done, &status, &its));
      if(done) break;
    }
    return tidy_cg(M, C, dx, z, sol, n, res);
  }

  if(bcg) {
    TV *r0 = base, *r = base + n, *sv = base + 2 * (size_t)n, *t = base + 3 * (size_t)n,
       *v = base + 4 * (size_t)n, *p = base + 5 * (size_t)n, *z = base + 6 * (size_t)n,
       *dx = base + 7 * (size_t)n;
    GDP_HIP(hipMemsetAsync(dx, 0, sizeof(TV) * n, s));  // bisinit zeroes the work array
    GDP_TRY(op_apply(A, sol, t, C, 0));
    hipLaunchKernelGGL(k_bs_init<TV>, g, blk, 0, s, n, rhs, t, r0, r, p, part);
    fin<FIN_BS_INIT>(C, part, nb, 0);
    const bool rp = M.kind != PREC_NONE;
    GraphCache gc;
    const int batch = 8;
    // five kernels per iteration (plus two preconditioner applications):
    // SpMV + <r0,v>, s, SpMV + <t,t>/<t,s>, update + |r|^2/<r,r0>, p
    constexpr int CGB = 1024;
    const int ncg = std::max(1, std::min(ctx().cu_count, (n + CGB - 1) / CGB));
    // partial-sum slots: <r0,v> (nrow), <t,t> and <t,s> (2 x nrow),
    // |r|^2 and <r,r0> (2 x ncg)
    const long rows = (long)part_rows(n);
    double *part_t = part + rows, *part_r = part + 3 * rows;
    auto body = [&]() -> int {
      for(int it = 0; it < batch; it++) {
        const TV *src = p;
        if(rp) {
          GDP_TRY(prec_apply(M, p, z, C, 1, n));
          src = z;
        }
        hipLaunchKernelGGL((k_op_dot<TA, TV, true>), grow, dim3(256), 0, s, A, src, r0, v, C, 1, part);
        hipLaunchKernelGGL((k_bs_s<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, r, v, sv, C, part, nrow,
                           it & 1);
        src = sv;
        if(rp) {
          GDP_TRY(prec_apply(M, sv, z, C, 1, n));
          src = z;
        }
        hipLaunchKernelGGL((k_op_dot2<TA, TV>), grow, dim3(256), 0, s, A, src, sv, t, C, part_t, nrow);
        hipLaunchKernelGGL((k_bs_update<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, p, sv, t, r0, r, dx, C,
                           part_t, nrow, part_r, ncg);
        hipLaunchKernelGGL((k_bs_p<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, r, v, p, C, part_r, ncg,
                           it & 1);
      }
      GDP_HIP(hipGetLastError());
      return GDP_OK;
    };
    GDP_TRY(read_ctl(C, &done, &status, &its));
    while(!done) {
      GDP_TRY(run_maybe_graph(opt.use_graphs, gc, body));
      GDP_TRY(read_ctl(C, &done, &status, &its));
    }
    return tidy_cg(M, C, dx, z, sol, n, res);
  }

  if(tfq) {
    TV *r0 = base, *w = base + n, *y = base + 2 * (size_t)n, *z = base + 3 * (size_t)n,
       *ay = base + 4 * (size_t)n, *az = base + 5 * (size_t)n, *v = base + 6 * (size_t)n,
       *d = base + 7 * (size_t)n, *tmp = base + 8 * (size_t)n, *dx = base + 9 * (size_t)n;
    // bisinit zeroes the work array: z, v, d and dx are read before written
    GDP_HIP(hipMemsetAsync(z, 0, sizeof(TV) * n, s));
    GDP_HIP(hipMemsetAsync(v, 0, sizeof(TV) * n, s));
    GDP_HIP(hipMemsetAsync(d, 0, sizeof(TV) * n, s));
    GDP_HIP(hipMemsetAsync(dx, 0, sizeof(TV) * n, s));
    GDP_TRY(op_apply(A, sol, az, C, 0));
    constexpr int CGB = 1024;
    const int ncg = std::max(1, std::min(ctx().cu_count, (n + CGB - 1) / CGB));
    // partial-sum slots: <r0,v> (nrow), |w|^2 after the second product
    // (nrow), |w|^2 and <r0,w> after update I (2 x ncg)
    const long rows = (long)part_rows(n);
    double *part_v = part, *part_w = part + rows, *part_u = part + 2 * rows;
    hipLaunchKernelGGL((k_tq_init<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, rhs, az, r0, w, C, part_u + ncg);
    const bool rp = M.kind != PREC_NONE;
    GraphCache gc;
    const int batch = 8;
    // per iteration: y (+ rho), SpMV + v + <r0,v>, z (+ alpha),
    // SpMV + w + |w|^2, update I (+ |w|^2, <r0,w>), update II
    auto body = [&]() -> int {
      for(int it = 0; it < batch; it++) {
        const int par = it & 1;
        hipLaunchKernelGGL((k_tq_y<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, w, z, y, C, part_u + ncg, ncg,
                           par);
        const TV *src = y;
        if(rp) {
          GDP_TRY(prec_apply(M, y, tmp, C, 1, n));
          src = tmp;
        }
        hipLaunchKernelGGL((k_tq_op_v<TA, TV>), grow, dim3(256), 0, s, A, src, az, r0, ay, v, C, part_v);
        hipLaunchKernelGGL((k_tq_z<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, y, v, z, C, part_v, nrow, par);
        src = z;
        if(rp) {
          GDP_TRY(prec_apply(M, z, tmp, C, 1, n));
          src = tmp;
        }
        hipLaunchKernelGGL((k_tq_op_w<TA, TV>), grow, dim3(256), 0, s, A, src, ay, az, w, C, part_w);
        hipLaunchKernelGGL((k_tq_upd1<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, y, az, r0, d, dx, w, C, part_w,
                           nrow, part_u, ncg);
        hipLaunchKernelGGL((k_tq_upd2<TV, CGB>), dim3(ncg), dim3(CGB), 0, s, n, z, d, dx, C, part_u, ncg);
      }
      GDP_HIP(hipGetLastError());
      return GDP_OK;
    };
    for(;;) {
      GDP_TRY(run_maybe_graph(opt.use_graphs, gc, body));
      GDP_TRY(read_ctl(C, &done, &status, &its));
      if(done) break;
    }
    return tidy_cg(M, C, dx, tmp, sol, n, res);
  }

  // ---- (F)GMRES ----
  TV *V = base;                        // m+1 basis vectors
  TV *Z = base + (size_t)(m + 1) * n;  // FGMRES: m vectors; GMRES: 1 temp
  GraphCache gc;
  GmArgs<TA, TV> ga{A, &M, rhs, sol, V, Z, C, part, n, m, opt.orth, flexible, two_stage};
  auto cycle = [&]() -> int { return gm_cycle<TA, TV>(ga); };
  for(int cyc = 0;; cyc++) {
    // the first two restart cycles launch their kernels directly: a solve
    // that ends within them does not pay for a capture it would replay once
    // at most
    GDP_TRY(run_maybe_graph(opt.use_graphs && cyc >= 2, gc, cycle));
    GDP_TRY(read_ctl(C, &done, &status, &its));
    if(done) break;
  }
  if(M.kind == PREC_ILU) GDP_TRY(M.ilu->check_timeout());
  KCtl tmp;
  GDP_TRY(d2h(&tmp, C, 1));
  res.its = tmp.its;
  res.status = tmp.status;
  res.res0 = tmp.fpar3;
  res.res_est = tmp.fpar6;
  return GDP_OK;
}

// ---- nested FP32 GMRES as a preconditioner ------------------------------------------------

namespace {

// Fresh control block for one inner solve: k steps, no tolerance stop; when
// the outer solve has already stopped (gate) the whole inner solve is skipped.
__global__ void __launch_bounds__(256) k_inner_begin(KCtl *Ci, const KCtl *outer, int gmode, int k)
{
  double *d = (double *)Ci;
  const int nd = (int)(sizeof(KCtl) / sizeof(double));
  for(int i = threadIdx.x; i < nd; i += 256) d[i] = 0.;
  __syncthreads();
  if(threadIdx.x == 0) {
    Ci->done = gated(outer, gmode) ? 1 : 0;
    Ci->maxits = k;
    Ci->m = k;
    Ci->rtol = 0.;
  }
}

__global__ void __launch_bounds__(B) k_d2f_gated(int n, const double *v, float *r, float *x, const KCtl *outer,
                                                 int gmode)
{
  if(gated(outer, gmode)) return;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) {
    r[j] = (float)v[j];
    x[j] = 0.f;
  }
}

__global__ void __launch_bounds__(B) k_f2d_gated(int n, const float *x, double *z, const KCtl *outer, int gmode)
{
  if(gated(outer, gmode)) return;
  for(long j = (long)blockIdx.x * B + threadIdx.x; j < n; j += (long)gridDim.x * B) z[j] = (double)x[j];
}

}  // namespace

InnerSolver::~InnerSolver()
{
  dfree(C);
  dfree(part);
  dfree(vec);
}

int InnerSolver::prepare()
{
  n = A.n;
  k = std::max(2, std::min(k, kMaxKrylov));
  if(!C) GDP_TRY(dalloc(&C, 1));
  if(!part) {
    const size_t rows = (size_t)std::max(1024L, ((long)n + 255) / 256);
    GDP_TRY(dalloc(&part, rows * (kMaxKrylov + 2)));
  }
  if(!vec) GDP_TRY(dalloc(&vec, (size_t)(k + 4) * std::max(1, n)));
  return GDP_OK;
}

int inner_apply(InnerSolver *S, const double *v, double *z, KCtl *outer, int gmode)
{
  if(!S || !S->vec) return set_error(GDP_E_STATE, "inner FP32 solver not prepared");
  hipStream_t s = ctx().stream;
  const int n = S->n, k = S->k;
  float *rhs = S->vec, *x = rhs + n, *V = x + n;  // V: k + 1 basis vectors, then one temporary
  hipLaunchKernelGGL(k_inner_begin, dim3(1), dim3(256), 0, s, S->C, outer, gmode, k);
  dim3 g(reduce_blocks(n));
  hipLaunchKernelGGL(k_d2f_gated, g, dim3(B), 0, s, n, v, rhs, x, outer, gmode);
  GmArgs<float, float> ga{S->A, &S->M, rhs, x, V, V + (size_t)(k + 1) * n, S->C, S->part, n, k, 1, false, false};
  GDP_TRY((gm_cycle<float, float>(ga)));
  hipLaunchKernelGGL(k_f2d_gated, g, dim3(B), 0, s, n, x, z, outer, gmode);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

template int krylov_solve<double, double>(const DevOp<double> &, const PrecOp<double> &,
                                          const double *, double *, const KrylovOptions &,
                                          KrylovResult &);
template int krylov_solve<float, float>(const DevOp<float> &, const PrecOp<float> &,
                                        const float *, float *, const KrylovOptions &,
                                        KrylovResult &);

}  // namespace gdp
