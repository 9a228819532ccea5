/*
 * gdp_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU oracle.
 *
 * Plain-C restatement of the reference's FP64 assembly + Sparskit solve.
 * Each function cites the reference code it follows and keeps that code's
 * order of floating-point operations, so that results can be compared bit
 * for bit with the reference's own Fortran kernels (oracle/_ref) and with
 * closed forms.  Compiled with -ffp-contract=off (no fused multiply-add).
 *
 * Never loaded by the product; see gdp_oracle.h for the conventions.
 */
#include "gdp_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ========================================================================
 * Sparskit kernels
 * ====================================================================== */

/* BLAS-1 ddot as vendored by the reference (contrib/Sparskit/blas1.f:52-100):
 * clean-up loop of n mod 5 terms, then a 5-way unrolled sum evaluated left to
 * right.  Sparskit's distdot (iters.f:3-10) is this ddot. */
static int g_dot_blocks = 0;

/* Tests only: with nblocks > 0 every dot product is summed in nblocks
 * contiguous blocks whose sums are then added in order -- the reference's
 * algorithm with a different rounding order (as a blocked device reduction
 * has), used to measure how far rounding alone moves a result.  0 restores
 * the reference's order. */
void orc_set_dot_blocks(int nblocks) { g_dot_blocks = nblocks > 0 ? nblocks : 0; }

/* CPU baselines only: with nthreads > 1 the element loop of orc_assemble
 * (bit-identical result), amux (row-parallel, bit-identical) and the vector
 * loops of CG and GMRES run on that many OpenMP threads, and dot products are
 * summed in 64 contiguous blocks added in order (deterministic, a different
 * rounding order from the reference's).  ILU factorisation and triangular
 * solves stay sequential, as in the reference.  1 (the default) is the
 * reference's single-threaded order everywhere. */
static int g_threads = 1;
void orc_set_threads(int nthreads) { g_threads = nthreads > 1 ? nthreads : 1; }
#define PFOR _Pragma("omp parallel for schedule(static) if(g_threads > 1 && n >= PAR_N) num_threads(g_threads)")
#define PAR_N 16384 /* vectors shorter than this stay on one thread */

double orc_ddot(int n, const double *x, const double *y)
{
  double t = 0.0;
  if(n <= 0) return 0.0;
  if(g_threads > 1 && g_dot_blocks == 0 && n >= PAR_N) {
    double part[64];
    const int per = (n + 63) / 64;
    PFOR
    for(int b = 0; b < 64; b++) {
      double s = 0.0;
      for(int i = b * per; i < n && i < (b + 1) * per; i++) s = s + x[i] * y[i];
      part[b] = s;
    }
    for(int b = 0; b < 64; b++) t = t + part[b];
    return t;
  }
  if(g_dot_blocks > 0) {
    const int nb = g_dot_blocks, per = (n + nb - 1) / nb;
    for(int b = 0; b < nb; b++) {
      double s = 0.0;
      for(int i = b * per; i < n && i < (b + 1) * per; i++) s = s + x[i] * y[i];
      t = t + s;
    }
    return t;
  }
  int m = n % 5;
  for(int i = 0; i < m; i++) t = t + x[i] * y[i];
  if(n < 5) return t;
  for(int i = m; i < n; i += 5)
    t = t + x[i] * y[i] + x[i + 1] * y[i + 1] + x[i + 2] * y[i + 2] +
        x[i + 3] * y[i + 3] + x[i + 4] * y[i + 4];
  return t;
}

/* amux, contrib/Sparskit/matvec.f:35-76: row-dot CSR product. */
void orc_amux(int n, const double *x, double *y, const int *ia, const int *ja,
              const double *a)
{
  PFOR
  for(int i = 0; i < n; i++) {
    double t = 0.0;
    for(int k = ia[i]; k < ia[i + 1]; k++) t = t + a[k] * x[ja[k]];
    y[i] = t;
  }
}

/* ilu0, contrib/Sparskit/ilut.F:1818-1941.  With HAVE_ILU_FLOAT (the
 * reference's default, CMakeLists.txt:552) alu is real*4 while the scalar
 * temporaries are real*8 (implicit real*8 (a-h,o-z)): the product
 * alu(j)*alu(jrow) is formed in single precision, the update
 * alu(jw) - tl*alu(jj) in double precision and rounded on store. */
#define ILU0_BODY(T, MULF)                                                     \
  int *iw = (int *)calloc((size_t)n + 1, sizeof(int));                         \
  int ju0 = n + 1;                                                             \
  jlu[0] = ju0;                                                                \
  for(int ii = 0; ii < n; ii++) {                                              \
    int js = ju0;                                                              \
    ju[ii] = ju0;                                                              \
    for(int j = ia[ii]; j < ia[ii + 1]; j++) {                                 \
      int jcol = ja[j];                                                        \
      if(jcol == ii) {                                                         \
        alu[ii] = (T)a[j];                                                     \
        iw[jcol] = ii + 1;                                                     \
        ju[ii] = ju0;                                                          \
      }                                                                        \
      else {                                                                   \
        alu[ju0] = (T)a[j];                                                    \
        jlu[ju0] = jcol;                                                       \
        iw[jcol] = ju0 + 1;                                                    \
        ju0++;                                                                 \
      }                                                                        \
    }                                                                          \
    jlu[ii + 1] = ju0;                                                         \
    int jf = ju0 - 1, jm = ju[ii] - 1;                                         \
    for(int j = js; j <= jm; j++) {                                            \
      int jrow = jlu[j];                                                       \
      double tl = MULF(alu[j], alu[jrow]);                                     \
      alu[j] = (T)tl;                                                          \
      for(int jj = ju[jrow]; jj < jlu[jrow + 1]; jj++) {                       \
        int jw = iw[jlu[jj]];                                                  \
        if(jw != 0)                                                            \
          alu[jw - 1] = (T)((double)alu[jw - 1] - tl * (double)alu[jj]);       \
      }                                                                        \
    }                                                                          \
    if(alu[ii] == 0) {                                                         \
      free(iw);                                                                \
      return ii + 1;                                                           \
    }                                                                          \
    alu[ii] = (T)(1.0 / (double)alu[ii]);                                      \
    iw[ii] = 0;                                                                \
    for(int i = js; i <= jf; i++) iw[jlu[i]] = 0;                              \
  }                                                                            \
  free(iw);                                                                    \
  return 0;

static inline double mul_f32(float a, float b)
{
  float p = a * b; /* real*4 * real*4 is a real*4 product */
  return (double)p;
}
static inline double mul_f64(double a, double b) { return a * b; }

int orc_ilu0_f32(int n, const int *ia, const int *ja, const double *a,
                 float *alu, int *jlu, int *ju)
{
  ILU0_BODY(float, mul_f32)
}

int orc_ilu0_f64(int n, const int *ia, const int *ja, const double *a,
                 double *alu, int *jlu, int *ju)
{
  ILU0_BODY(double, mul_f64)
}

/* qsplit, ilut.F:2403-2460: partial quick-split so that the ncut entries of
 * largest magnitude come first (their order is the algorithm's). 1-based. */
static void qsplit1(double *a, int *ind, int n, int ncut)
{
  int first = 1, last = n;
  if(ncut < first || ncut > last) return;
  for(;;) {
    int mid = first;
    double abskey = fabs(a[mid]);
    for(int j = first + 1; j <= last; j++) {
      if(fabs(a[j]) > abskey) {
        mid = mid + 1;
        double tmp = a[mid];
        int itmp = ind[mid];
        a[mid] = a[j];
        ind[mid] = ind[j];
        a[j] = tmp;
        ind[j] = itmp;
      }
    }
    double tmp = a[mid];
    a[mid] = a[first];
    a[first] = tmp;
    int itmp = ind[mid];
    ind[mid] = ind[first];
    ind[first] = itmp;
    if(mid == ncut) return;
    if(mid > ncut) last = mid - 1;
    else first = mid + 1;
  }
}

/* ilut (ilut.F:52-396, pivot = 0) and ilutp (:397-806, pivot = 1, mbloc = n
 * as solve_matrix passes it), with `scalar` = T for the factor storage: the
 * products read alu in T and promote, stores round to T.  Written with the
 * Fortran's 1-based indices; the MSR outputs are converted to 0-based.
 * ilut's zero-pivot replacement uses the single-precision literal 0.0001,
 * ilutp's the double 1.0D-4.  ilutp leaves ja (0-based here) permuted and
 * iperm (2n, 0-based values) with iperm[0..n) = old numbers, iperm[n..2n) =
 * new numbers. */
#define ILUT_BODY(T)                                                                          \
  if(lfil < 0) return -4;                                                                     \
  int *IA = (int *)malloc(sizeof(int) * (size_t)(n + 2));                                     \
  int *JW = (int *)calloc((size_t)2 * n + 2, sizeof(int));                                    \
  int *IP = (int *)malloc(sizeof(int) * ((size_t)2 * n + 2));                                 \
  double *W = (double *)calloc((size_t)n + 2, sizeof(double));                                \
  T *ALU = (T *)calloc((size_t)iwk + 2, sizeof(T));                                           \
  int *JLU = (int *)calloc((size_t)iwk + 2, sizeof(int));                                     \
  int *JU = (int *)calloc((size_t)n + 2, sizeof(int));                                        \
  int ierr = 0;                                                                               \
  for(int i = 0; i <= n; i++) IA[i + 1] = ia[i] + 1;                                          \
  int ju0 = n + 2;                                                                            \
  JLU[1] = ju0;                                                                               \
  for(int j = 1; j <= n; j++) {                                                               \
    JW[n + j] = 0;                                                                            \
    IP[j] = j;                                                                                \
    IP[n + j] = j;                                                                            \
  }                                                                                           \
  for(int ii = 1; ii <= n && !ierr; ii++) {                                                   \
    int j1 = IA[ii], j2 = IA[ii + 1] - 1;                                                     \
    double tnorm = 0.0;                                                                       \
    for(int k = j1; k <= j2; k++) tnorm = tnorm + fabs(a[k - 1]);                             \
    if(tnorm == 0.0) {                                                                        \
      ierr = -5;                                                                              \
      break;                                                                                  \
    }                                                                                         \
    tnorm = tnorm / (double)(j2 - j1 + 1);                                                    \
    int lenu = 1, lenl = 0;                                                                   \
    JW[ii] = ii;                                                                              \
    W[ii] = 0.0;                                                                              \
    JW[n + ii] = ii;                                                                          \
    for(int j = j1; j <= j2; j++) {                                                           \
      int k = IP[n + (ja[j - 1] + 1)];                                                        \
      double t = a[j - 1];                                                                    \
      if(k < ii) {                                                                            \
        lenl = lenl + 1;                                                                      \
        JW[lenl] = k;                                                                         \
        W[lenl] = t;                                                                          \
        JW[n + k] = lenl;                                                                     \
      }                                                                                       \
      else if(k == ii)                                                                        \
        W[ii] = t;                                                                            \
      else {                                                                                  \
        lenu = lenu + 1;                                                                      \
        int jpos = ii + lenu - 1;                                                             \
        JW[jpos] = k;                                                                         \
        W[jpos] = t;                                                                          \
        JW[n + k] = jpos;                                                                     \
      }                                                                                       \
    }                                                                                         \
    int jj = 0, len = 0;                                                                      \
    for(;;) {                                                                                 \
      jj = jj + 1;                                                                            \
      if(jj > lenl) break;                                                                    \
      int jrow = JW[jj], k = jj;                                                              \
      for(int j = jj + 1; j <= lenl; j++)                                                     \
        if(JW[j] < jrow) {                                                                    \
          jrow = JW[j];                                                                       \
          k = j;                                                                              \
        }                                                                                     \
      if(k != jj) {                                                                           \
        int j = JW[jj];                                                                       \
        JW[jj] = JW[k];                                                                       \
        JW[k] = j;                                                                            \
        JW[n + jrow] = jj;                                                                    \
        JW[n + j] = k;                                                                        \
        double sw = W[jj];                                                                    \
        W[jj] = W[k];                                                                         \
        W[k] = sw;                                                                            \
      }                                                                                       \
      JW[n + jrow] = 0;                                                                       \
      double fact = W[jj] * (double)ALU[jrow];                                                \
      if(fabs(fact) <= droptol) continue;                                                     \
      for(int kk = JU[jrow]; kk <= JLU[jrow + 1] - 1; kk++) {                                 \
        double sv = fact * (double)ALU[kk];                                                   \
        int j = IP[n + JLU[kk]];                                                              \
        int jpos = JW[n + j];                                                                 \
        if(j >= ii) {                                                                         \
          if(jpos == 0) {                                                                     \
            lenu = lenu + 1;                                                                  \
            if(lenu > n) {                                                                    \
              ierr = -1;                                                                      \
              break;                                                                          \
            }                                                                                 \
            int i = ii + lenu - 1;                                                            \
            JW[i] = j;                                                                        \
            JW[n + j] = i;                                                                    \
            W[i] = -sv;                                                                       \
          }                                                                                   \
          else                                                                                \
            W[jpos] = W[jpos] - sv;                                                           \
        }                                                                                     \
        else {                                                                                \
          if(jpos == 0) {                                                                     \
            lenl = lenl + 1;                                                                  \
            if(lenl > n) {                                                                    \
              ierr = -1;                                                                      \
              break;                                                                          \
            }                                                                                 \
            JW[lenl] = j;                                                                     \
            JW[n + j] = lenl;                                                                 \
            W[lenl] = -sv;                                                                    \
          }                                                                                   \
          else                                                                                \
            W[jpos] = W[jpos] - sv;                                                           \
        }                                                                                     \
      }                                                                                       \
      if(ierr) break;                                                                         \
      len = len + 1;                                                                          \
      W[len] = fact;                                                                          \
      JW[len] = jrow;                                                                         \
    }                                                                                         \
    if(ierr) break;                                                                           \
    for(int k = 1; k <= lenu; k++) JW[n + JW[ii + k - 1]] = 0;                                \
    lenl = len;                                                                               \
    len = lenl < lfil ? lenl : lfil;                                                          \
    qsplit1(W, JW, lenl, len);                                                                \
    for(int k = 1; k <= len; k++) {                                                           \
      if(ju0 > iwk) {                                                                         \
        ierr = -2;                                                                            \
        break;                                                                                \
      }                                                                                       \
      ALU[ju0] = (T)W[k];                                                                     \
      JLU[ju0] = IP[JW[k]];                                                                   \
      ju0 = ju0 + 1;                                                                          \
    }                                                                                         \
    if(ierr) break;                                                                           \
    JU[ii] = ju0;                                                                             \
    len = 0;                                                                                  \
    for(int k = 1; k <= lenu - 1; k++)                                                        \
      if(fabs(W[ii + k]) > droptol * tnorm) {                                                 \
        len = len + 1;                                                                        \
        W[ii + len] = W[ii + k];                                                              \
        JW[ii + len] = JW[ii + k];                                                            \
      }                                                                                       \
    lenu = len + 1;                                                                           \
    len = lenu < lfil ? lenu : lfil;                                                          \
    qsplit1(W + ii, JW + ii, lenu - 1, len);                                                  \
    if(pivot) {                                                                               \
      int imax = ii;                                                                          \
      double xmax = fabs(W[imax]), xmax0 = xmax;                                              \
      int icut = ii - 1 + n - (ii - 1) % n;                                                   \
      for(int k = ii + 1; k <= ii + len - 1; k++) {                                           \
        double t = fabs(W[k]);                                                                \
        if(t > xmax && t * permtol > xmax0 && JW[k] <= icut) {                                \
          imax = k;                                                                           \
          xmax = t;                                                                           \
        }                                                                                     \
      }                                                                                       \
      double tmp = W[ii];                                                                     \
      W[ii] = W[imax];                                                                        \
      W[imax] = tmp;                                                                          \
      int j = JW[imax];                                                                       \
      int i = IP[ii];                                                                         \
      IP[ii] = IP[j];                                                                         \
      IP[j] = i;                                                                              \
      IP[n + IP[ii]] = ii;                                                                    \
      IP[n + IP[j]] = j;                                                                      \
    }                                                                                         \
    if(len + ju0 > iwk) {                                                                     \
      ierr = -3;                                                                              \
      break;                                                                                  \
    }                                                                                         \
    for(int k = ii + 1; k <= ii + len - 1; k++) {                                             \
      JLU[ju0] = IP[JW[k]];                                                                   \
      ALU[ju0] = (T)W[k];                                                                     \
      ju0 = ju0 + 1;                                                                          \
    }                                                                                         \
    if(W[ii] == 0.0) W[ii] = (pivot ? (1.0e-4 + droptol) : ((double)0.0001f + droptol)) * tnorm; \
    ALU[ii] = (T)(1.0 / W[ii]);                                                               \
    JLU[ii + 1] = ju0;                                                                        \
  }                                                                                           \
  if(!ierr) {                                                                                 \
    if(pivot) {                                                                               \
      for(int k = JLU[1]; k <= JLU[n + 1] - 1; k++) JLU[k] = IP[n + JLU[k]];                  \
      for(int k = 0; k < ia[n]; k++) ja[k] = IP[n + ja[k] + 1] - 1;                           \
    }                                                                                         \
    for(int k = 1; k <= ju0; k++) {                                                           \
      alu[k - 1] = ALU[k];                                                                    \
      jlu[k - 1] = JLU[k] - 1;                                                                \
    }                                                                                         \
    for(int i = 1; i <= n; i++) ju[i - 1] = JU[i] - 1;                                        \
    if(iperm)                                                                                 \
      for(int i = 1; i <= 2 * n; i++) iperm[i - 1] = IP[i] - 1;                               \
  }                                                                                           \
  free(IA);                                                                                   \
  free(JW);                                                                                   \
  free(IP);                                                                                   \
  free(W);                                                                                    \
  free(ALU);                                                                                  \
  free(JLU);                                                                                  \
  free(JU);                                                                                   \
  return ierr;

int orc_ilut_f32(int n, const int *ia, int *ja, const double *a, int lfil, double droptol,
                 double permtol, int pivot, int iwk, float *alu, int *jlu, int *ju, int *iperm)
{
  ILUT_BODY(float)
}

int orc_ilut_f64(int n, const int *ia, int *ja, const double *a, int lfil, double droptol,
                 double permtol, int pivot, int iwk, double *alu, int *jlu, int *ju, int *iperm)
{
  ILUT_BODY(double)
}

/* lusol, contrib/Sparskit/ilut.F:2297-2347: unit-L forward sweep, then U
 * backward sweep with the stored inverted diagonal.  In place allowed. */
void orc_lusol_f32(int n, const double *y, double *x, const float *alu,
                   const int *jlu, const int *ju)
{
  for(int i = 0; i < n; i++) {
    x[i] = y[i];
    for(int k = jlu[i]; k < ju[i]; k++) x[i] = x[i] - (double)alu[k] * x[jlu[k]];
  }
  for(int i = n - 1; i >= 0; i--) {
    for(int k = ju[i]; k < jlu[i + 1]; k++)
      x[i] = x[i] - (double)alu[k] * x[jlu[k]];
    x[i] = (double)alu[i] * x[i];
  }
}

void orc_lusol_f64(int n, const double *y, double *x, const double *alu,
                   const int *jlu, const int *ju)
{
  for(int i = 0; i < n; i++) {
    x[i] = y[i];
    for(int k = jlu[i]; k < ju[i]; k++) x[i] = x[i] - alu[k] * x[jlu[k]];
  }
  for(int i = n - 1; i >= 0; i--) {
    for(int k = ju[i]; k < jlu[i + 1]; k++) x[i] = x[i] - alu[k] * x[jlu[k]];
    x[i] = alu[i] * x[i];
  }
}

/* ---- Krylov helpers ---------------------------------------------------- */

/* mgsro, iters.f:3475-3597, with full = .false.: modified Gram-Schmidt of
 * vector `ind` against vectors 0..ind-1 with selective re-orthogonalisation
 * (threshold 0.98 of the running norm estimate).  hh[0..ind] receives the
 * Hessenberg column. */
static int mgsro(int n, double *V, int ind, double *hh)
{
  double *v = V + (size_t)ind * n;
  double nrm0 = orc_ddot(n, v, v);
  double thr = nrm0 * 0.98;
  if(nrm0 <= 0.0) return -1;
  if(!(nrm0 > 0.0 && 1.0 / nrm0 > 0.0)) return -2;
  for(int i = 0; i < ind; i++) {
    const double *vi = V + (size_t)i * n;
    double fct = orc_ddot(n, v, vi);
    hh[i] = fct;
    PFOR
    for(int k = 0; k < n; k++) v[k] = v[k] - fct * vi[k];
    if(fct * fct > thr) {
      fct = orc_ddot(n, v, vi);
      hh[i] = hh[i] + fct;
      PFOR
      for(int k = 0; k < n; k++) v[k] = v[k] - fct * vi[k];
    }
    nrm0 = nrm0 - hh[i] * hh[i];
    if(nrm0 < 0.0) nrm0 = 0.0;
    thr = nrm0 * 0.98;
  }
  double nrm1 = sqrt(orc_ddot(n, v, v));
  hh[ind] = nrm1;
  if(nrm1 <= 0.0) return -3;
  double fct = 1.0 / nrm1;
  PFOR
  for(int k = 0; k < n; k++) v[k] = v[k] * fct;
  return 0;
}

/* givens, iters.f:3235-3272 (Golub-van Loan): rotation taking (x, y) to
 * (r, 0); x receives |r| computed as |x*y| after the in-place rescaling. */
static void givens(double *x, double *y, double *c, double *s)
{
  if(*x == 0.0 && *y == 0.0) {
    *c = 1.0;
    *s = 0.0;
  }
  else if(fabs(*y) > fabs(*x)) {
    double t = *x / *y;
    *x = sqrt(1.0 + t * t);
    *s = copysign(1.0 / *x, *y);
    *c = t * *s;
  }
  else if(fabs(*y) <= fabs(*x)) {
    double t = *y / *x;
    *y = sqrt(1.0 + t * t);
    *c = copysign(1.0 / *y, *x);
    *s = t * *c;
  }
  else {
    *x = 0.0;
    *y = 0.0;
    *c = 1.0;
    *s = 0.0;
  }
  *x = fabs(*x * *y);
}

/* Shared restarted (F)GMRES body: iters.f:1912-2217 (gmres) and
 * iters.f:2522-2807 (fgmres), with ipar(2) = 2 (right preconditioning when a
 * preconditioner is given), ipar(3) = 1 (stop on the estimated residual
 * relative to the first residual), fpar(2) = 0. */
static int gmres_body(int flexible, int n, const double *rhs, double *sol,
                      int m_in, int maxits, double rtol, orc_op matvec,
                      orc_op precond, void *ctx, orc_info *info)
{
  int m = m_in <= 1 ? 15 : m_in;
  int rp = precond != NULL;
  size_t N = (size_t)n;
  double *V = (double *)calloc(N * (m + 1), sizeof(double));
  double *Z = (double *)calloc(flexible ? N * m : N, sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * (m + 1), sizeof(double));
  double *cs = (double *)calloc((size_t)m + 1, sizeof(double));
  double *sn = (double *)calloc((size_t)m + 1, sizeof(double));
  double *g = (double *)calloc((size_t)m + 2, sizeof(double));
  int its = 0, status = 1, ierr12 = 0;
  double fpar3 = 0, fpar4 = 0, fpar5 = 0, fpar6 = 0;

  for(;;) { /* label 100: (re)start */
    matvec(ctx, sol, V);
    its++;
    PFOR
    for(int i = 0; i < n; i++) V[i] = rhs[i] - V[i];
    double alpha = sqrt(orc_ddot(n, V, V));
    if(its == 1) {
      fpar4 = rtol * alpha + 0.0;
      fpar3 = alpha;
    }
    fpar5 = alpha;
    g[0] = alpha;
    if(alpha <= fpar4) {
      status = 0;
      fpar6 = alpha;
      goto done;
    }
    alpha = 1.0 / alpha;
    PFOR
    for(int i = 0; i < n; i++) V[i] = alpha * V[i];

    int k = 0;
    for(;;) { /* label 110: one Arnoldi step */
      k++;
      double *vk = V + (size_t)(k - 1) * N, *vn = V + (size_t)k * N;
      if(flexible) {
        double *zk = Z + (size_t)(k - 1) * N;
        if(rp) precond(ctx, vk, zk);
        else memcpy(zk, vk, N * sizeof(double));
        matvec(ctx, rp ? zk : vk, vn);
      }
      else if(rp) {
        precond(ctx, vk, Z);
        matvec(ctx, Z, vn);
      }
      else
        matvec(ctx, vk, vn);
      its++;
      double *hc = H + (size_t)(k - 1) * (m + 1);
      ierr12 = mgsro(n, V, k, hc);
      if(ierr12 < 0) break;
      for(int i = 0; i < k - 1; i++) {
        double a0 = hc[i];
        hc[i] = cs[i] * a0 + sn[i] * hc[i + 1];
        hc[i + 1] = cs[i] * hc[i + 1] - sn[i] * a0;
      }
      givens(&hc[k - 1], &hc[k], &cs[k - 1], &sn[k - 1]);
      alpha = -sn[k - 1] * g[k - 1];
      g[k - 1] = cs[k - 1] * g[k - 1];
      g[k] = alpha;
      alpha = fabs(alpha);
      fpar5 = alpha;
      if(k < m && !(alpha <= fpar4) && (maxits <= 0 || its < maxits)) continue;
      break;
    }
    /* label 200: back substitution on the rotated Hessenberg system */
    while(H[(size_t)(k - 1) * (m + 1) + (k - 1)] == 0.0) {
      k--;
      if(k <= 0) {
        status = -3;
        goto done;
      }
    }
    g[k - 1] = g[k - 1] / H[(size_t)(k - 1) * (m + 1) + (k - 1)];
    for(int i = k - 1; i >= 1; i--) {
      const double *hcol = H + (size_t)i * (m + 1); /* column i+1 (1-based) */
      for(int ii = 0; ii < i; ii++) g[ii] = g[ii] - g[i] * hcol[ii];
      g[i - 1] = g[i - 1] / H[(size_t)(i - 1) * (m + 1) + (i - 1)];
    }
    if(flexible) {
      for(int i = 0; i < k; i++) {
        const double *zi = Z + (size_t)i * N;
        PFOR
        for(int ii = 0; ii < n; ii++) sol[ii] = sol[ii] + g[i] * zi[ii];
      }
    }
    else {
      PFOR
      for(int ii = 0; ii < n; ii++) V[ii] = V[ii] * g[0];
      for(int i = 1; i < k; i++) {
        const double *vi = V + (size_t)i * N;
        PFOR
        for(int ii = 0; ii < n; ii++) V[ii] = V[ii] + g[i] * vi[ii];
      }
      if(rp) {
        precond(ctx, V, Z);
        PFOR
        for(int i = 0; i < n; i++) sol[i] = sol[i] + Z[i];
      }
      else {
        PFOR
        for(int i = 0; i < n; i++) sol[i] = sol[i] + V[i];
      }
    }
    fpar6 = fpar5;
    if(ierr12 != 0) {
      status = -3;
      goto done;
    }
    if((its < maxits || maxits <= 0) && fpar6 > fpar4) continue;
    if(fpar6 <= fpar4) status = 0;
    else if(its >= maxits && maxits > 0) status = -1;
    else status = -10;
    break;
  }
done:
  info->its = its;
  info->status = status;
  info->res_est = fpar6;
  info->res0 = fpar3;
  free(V);
  free(Z);
  free(H);
  free(cs);
  free(sn);
  free(g);
  return status;
}

int orc_gmres(int n, const double *rhs, double *sol, int m, int maxits,
              double rtol, orc_op matvec, orc_op precond, void *ctx,
              orc_info *info)
{
  return gmres_body(0, n, rhs, sol, m, maxits, rtol, matvec, precond, ctx, info);
}

int orc_fgmres(int n, const double *rhs, double *sol, int m, int maxits,
               double rtol, orc_op matvec, orc_op precond, void *ctx,
               orc_info *info)
{
  return gmres_body(1, n, rhs, sol, m, maxits, rtol, matvec, precond, ctx, info);
}

/* brkdn, iters.f:3368-3401: zero -> -3, abnormal -> -9, else 0. */
static int brkdn(double alpha)
{
  if(alpha > 0.0) return (1.0 / alpha > 0.0) ? 0 : -9;
  if(alpha < 0.0) return (1.0 / alpha < 0.0) ? 0 : -9;
  if(alpha == 0.0) return -3;
  return -9;
}

/* cg, iters.f:379-562 with stopbis (:3275-3330) and tidycg (:3333-3365).
 * Sparskit's CG is not PCG: with a right preconditioner it iterates on
 * A*M^-1 and maps the accumulated step back through M^-1 at the end. */
int orc_cg(int n, const double *rhs, double *sol, int maxits, double rtol,
           orc_op matvec, orc_op precond, void *ctx, orc_info *info)
{
  size_t N = (size_t)n;
  double *r = (double *)calloc(N, sizeof(double));
  double *p = (double *)calloc(N, sizeof(double));
  double *ap = (double *)calloc(N, sizeof(double));
  double *z = (double *)calloc(N, sizeof(double));
  double *dx = (double *)calloc(N, sizeof(double));
  int rp = precond != NULL;
  int its = 0, status = 1, ipar11 = 0, ipar12 = 0;
  double fpar3, fpar4, fpar5, fpar6 = 0.0, fpar7;

  matvec(ctx, sol, ap);
  its++;
  for(int i = 0; i < n; i++) p[i] = rhs[i] - ap[i];
  for(int i = 0; i < n; i++) r[i] = p[i];
  fpar7 = orc_ddot(n, r, r);
  fpar3 = sqrt(fpar7);
  fpar5 = fpar3;
  fpar4 = rtol * fpar3 + 0.0;
  for(;;) {
    if(rp) {
      precond(ctx, p, z);
      matvec(ctx, z, ap);
    }
    else
      matvec(ctx, p, ap);
    its++;
    double alpha = orc_ddot(n, p, ap);
    int bk = brkdn(alpha);
    if(bk) {
      status = bk;
      break;
    }
    alpha = fpar7 / alpha;
    PFOR
    for(int i = 0; i < n; i++) {
      dx[i] = dx[i] + alpha * p[i];
      r[i] = r[i] - alpha * ap[i];
    }
    /* stopbis */
    int stop = (ipar11 == 1);
    if(maxits > 0 && its >= maxits) {
      status = -1;
      stop = 1;
    }
    if(!stop) {
      fpar5 = sqrt(orc_ddot(n, r, r));
      fpar6 = fpar5;
      if(fpar6 > fpar4) ipar11 = 0;
      else {
        ipar11 = 1;
        stop = 1;
      }
    }
    if(stop) break;
    alpha = fpar5 * fpar5 / fpar7;
    fpar7 = fpar5 * fpar5;
    PFOR
    for(int i = 0; i < n; i++) p[i] = r[i] + alpha * p[i];
  }
  const double *delta = dx;
  if(rp) {
    if(status < 0) ipar12 = status;
    status = 5;
    precond(ctx, dx, z);
    delta = z;
  }
  if(ipar12 != 0) status = ipar12;
  else if(status > 0) {
    if(fpar6 <= fpar4) status = 0;
    else if(its >= maxits && maxits > 0) status = -1;
    else status = -10;
  }
  for(int i = 0; i < n; i++) sol[i] = sol[i] + delta[i];
  info->its = its;
  info->status = status;
  info->res_est = fpar6;
  info->res0 = fpar3;
  free(r);
  free(p);
  free(ap);
  free(z);
  free(dx);
  return status;
}

/* bcgstab, iters.f:1032-1306 with stopbis (:3275-3330), tidycg (:3333-3365)
 * and brkdn (:3368-3401), for ipar(3) = 1.  Work columns of the reference:
 * r0 = w1, r = w2, s = w3, t = w4, v = w5, p = w6, z = w7 (M^-1 input of
 * the next product), dx = w8.  Right preconditioning iterates on A*M^-1 and
 * maps the accumulated step back through M^-1 at the end, as in orc_cg.
 * `its` counts matrix-vector products (ipar(7)): 1 + 2 per iteration. */
int orc_bcgstab(int n, const double *rhs, double *sol, int maxits, double rtol,
                orc_op matvec, orc_op precond, void *ctx, orc_info *info)
{
  size_t N = (size_t)n;
  double *r0 = (double *)calloc(N, sizeof(double));
  double *r = (double *)calloc(N, sizeof(double));
  double *s = (double *)calloc(N, sizeof(double));
  double *t = (double *)calloc(N, sizeof(double));
  double *v = (double *)calloc(N, sizeof(double));
  double *p = (double *)calloc(N, sizeof(double));
  double *z = (double *)calloc(N, sizeof(double));
  double *dx = (double *)calloc(N, sizeof(double));
  int rp = precond != NULL;
  int its = 0, status = 1, ipar12 = 0;
  double fpar3, fpar4, fpar5, fpar6, fpar7, alpha = 0.0, omega = 0.0;

  /* labels 10/20: r0 = rhs - A sol, r = p = r0 */
  matvec(ctx, sol, t);
  its++;
  for(int i = 0; i < n; i++) r0[i] = rhs[i] - t[i];
  for(int i = 0; i < n; i++) {
    r[i] = r0[i];
    p[i] = r0[i];
  }
  fpar7 = orc_ddot(n, r0, r0);
  fpar5 = sqrt(fpar7);
  fpar3 = fpar5;
  fpar4 = rtol * fpar3 + 0.0;
  fpar6 = fpar5;
  if(!(fpar5 <= fpar4)) {
    for(;;) {
      /* step (1): v = A M^-1 p */
      if(rp) {
        precond(ctx, p, z);
        matvec(ctx, z, v);
      }
      else
        matvec(ctx, p, v);
      its++;
      /* step (2) */
      double den = orc_ddot(n, r0, v);
      int bk = brkdn(den);
      if(bk) {
        status = bk;
        break;
      }
      alpha = fpar7 / den;
      /* step (3) */
      for(int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
      /* step (4): t = A M^-1 s */
      if(rp) {
        precond(ctx, s, z);
        matvec(ctx, z, t);
      }
      else
        matvec(ctx, s, t);
      its++;
      /* step (5) */
      double tt = orc_ddot(n, t, t);
      bk = brkdn(tt);
      if(bk) {
        status = bk;
        break;
      }
      omega = orc_ddot(n, t, s) / tt;
      bk = brkdn(omega);
      if(bk) {
        status = bk;
        break;
      }
      /* steps (6), (7) */
      for(int i = 0; i < n; i++) {
        z[i] = alpha * p[i] + omega * s[i];
        dx[i] = dx[i] + z[i];
        r[i] = s[i] - omega * t[i];
      }
      /* stopbis */
      if(maxits > 0 && its >= maxits) {
        status = -1;
        break;
      }
      fpar5 = sqrt(orc_ddot(n, r, r));
      fpar6 = fpar5;
      if(!(fpar6 > fpar4)) break;
      /* step (8) */
      double rho = fpar7;
      fpar7 = orc_ddot(n, r, r0);
      double beta = fpar7 * alpha / (omega * rho);
      for(int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
      bk = brkdn(fpar7);
      if(bk) {
        status = bk;
        break;
      }
    }
  }
  /* label 900 / 110 */
  const double *delta = dx;
  if(rp) {
    if(status < 0) ipar12 = status;
    status = 5;
    precond(ctx, dx, z);
    delta = z;
  }
  if(ipar12 != 0) status = ipar12;
  else if(status > 0) {
    if(fpar6 <= fpar4) status = 0;
    else if(its >= maxits && maxits > 0) status = -1;
    else status = -10;
  }
  for(int i = 0; i < n; i++) sol[i] = sol[i] + delta[i];
  info->its = its;
  info->status = status;
  info->res_est = fpar6;
  info->res0 = fpar3;
  free(r0);
  free(r);
  free(s);
  free(t);
  free(v);
  free(p);
  free(z);
  free(dx);
  return status;
}

/* tfqmr, iters.f:1307-1584, for ipar(3) = 1, with tidycg (:3333-3365) and
 * brkdn (:3368-3401).  Work columns: r0 = w2, w = w3, y = w4, z = w5,
 * ay = w6 = A M^-1 y, az = w7 = A M^-1 z (A sol at start), v = w8, d = w9,
 * M^-1 input/output w10, dx = w11; bisinit zeroes all of them.  There is no
 * test before the first iteration: a zero residual breaks down (rho = 0).
 * The residual estimate is tau (fpar(5) = tao, the "approximation" branch). */
int orc_tfqmr(int n, const double *rhs, double *sol, int maxits, double rtol,
              orc_op matvec, orc_op precond, void *ctx, orc_info *info)
{
  size_t N = (size_t)n;
  double *r0 = (double *)calloc(N, sizeof(double));
  double *w = (double *)calloc(N, sizeof(double));
  double *y = (double *)calloc(N, sizeof(double));
  double *z = (double *)calloc(N, sizeof(double));
  double *ay = (double *)calloc(N, sizeof(double));
  double *az = (double *)calloc(N, sizeof(double));
  double *v = (double *)calloc(N, sizeof(double));
  double *d = (double *)calloc(N, sizeof(double));
  double *tmp = (double *)calloc(N, sizeof(double));
  double *dx = (double *)calloc(N, sizeof(double));
  int rp = precond != NULL;
  int its = 0, status = 1, ipar12 = 0;
  double fpar3, fpar4, fpar5, fpar6 = 0.0;
  double eta, sigma, theta, te, alpha, rho, tao;

  /* labels 10/20 */
  matvec(ctx, sol, az);
  its++;
  for(int i = 0; i < n; i++) {
    r0[i] = rhs[i] - az[i];
    d[i] = 0.0;
  }
  for(int i = 0; i < n; i++) w[i] = r0[i];
  fpar5 = sqrt(orc_ddot(n, r0, r0));
  fpar3 = fpar5;
  tao = fpar5;
  fpar4 = rtol * tao + 0.0;
  te = 0.0;
  rho = 0.0;
  for(;;) {
    /* label 30 */
    sigma = rho;
    rho = orc_ddot(n, r0, w);
    int bk = brkdn(rho);
    if(bk) {
      status = bk;
      break;
    }
    alpha = (its == 1) ? 0.0 : rho / sigma;
    for(int i = 0; i < n; i++) y[i] = w[i] + alpha * z[i];
    if(rp) {
      precond(ctx, y, tmp);
      matvec(ctx, tmp, ay);
    }
    else
      matvec(ctx, y, ay);
    /* label 60 */
    its++;
    for(int i = 0; i < n; i++) v[i] = ay[i] + alpha * (az[i] + alpha * v[i]);
    sigma = orc_ddot(n, r0, v);
    bk = brkdn(sigma);
    if(bk) {
      status = bk;
      break;
    }
    alpha = rho / sigma;
    for(int i = 0; i < n; i++) z[i] = y[i] - alpha * v[i];
    if(rp) {
      precond(ctx, z, tmp);
      matvec(ctx, tmp, az);
    }
    else
      matvec(ctx, z, az);
    /* label 90 */
    its++;
    for(int i = 0; i < n; i++) w[i] = w[i] - alpha * ay[i];
    /* update I */
    theta = orc_ddot(n, w, w) / (tao * tao);
    sigma = 1.0 / (1.0 + theta);
    tao = tao * sqrt(sigma * theta);
    bk = brkdn(tao);
    if(bk) {
      status = bk;
      break;
    }
    eta = sigma * alpha;
    sigma = te / alpha;
    te = theta * eta;
    for(int i = 0; i < n; i++) {
      d[i] = y[i] + sigma * d[i];
      dx[i] = dx[i] + eta * d[i];
      w[i] = w[i] - alpha * az[i];
    }
    /* update II */
    theta = orc_ddot(n, w, w) / (tao * tao);
    sigma = 1.0 / (1.0 + theta);
    tao = tao * sqrt(sigma * theta);
    bk = brkdn(tao);
    if(bk) {
      status = bk;
      break;
    }
    eta = sigma * alpha;
    sigma = te / alpha;
    te = theta * eta;
    for(int i = 0; i < n; i++) {
      d[i] = z[i] + sigma * d[i];
      dx[i] = dx[i] + eta * d[i];
    }
    fpar5 = tao;
    fpar6 = fpar5;
    if(!(fpar6 > fpar4 && (its < maxits || maxits <= 0))) break;
  }
  /* label 900 / 110 */
  const double *delta = dx;
  if(rp) {
    if(status < 0) ipar12 = status;
    status = 5;
    precond(ctx, dx, tmp);
    delta = tmp;
  }
  if(ipar12 != 0) status = ipar12;
  else if(status > 0) {
    if(fpar6 <= fpar4) status = 0;
    else if(its >= maxits && maxits > 0) status = -1;
    else status = -10;
  }
  for(int i = 0; i < n; i++) sol[i] = sol[i] + delta[i];
  info->its = its;
  info->status = status;
  info->res_est = fpar6;
  info->res0 = fpar3;
  (void)fpar5;
  free(r0);
  free(w);
  free(y);
  free(z);
  free(ay);
  free(az);
  free(v);
  free(d);
  free(tmp);
  free(dx);
  return status;
}

/* ---- RCMK (cmkreord.f:3-62, reordering.f:15-391, unary.f:1292-1508) ------ */

/* BFS with a single start node and natural traversal order for further
 * components (iperm(1) = 0), reordering.f:15-156 + add_lvst :279-300.
 * riord/levels are 0-based here: levels[l] = first position of level l. */
static int bfs_levels(int n, const int *ia, const int *ja, int init, int *mask,
                      int *riord, int *levels)
{
  int nlev = 0, istart = 0, iend = 1, ii = 0;
  riord[0] = init;
  mask[init] = 0;
  for(;;) {
    levels[nlev++] = istart;
    int nod = iend;
    for(int ir = istart; ir < iend; ir++) {
      int i = riord[ir];
      for(int k = ia[i]; k < ia[i + 1]; k++) {
        int j = ja[k];
        if(mask[j] == 1) {
          mask[j] = 0;
          riord[nod++] = j;
        }
      }
    }
    istart = iend;
    iend = nod;
    if(istart < iend) continue;
    while(ii < n && mask[ii] != 1) ii++;
    if(ii >= n) break;
    istart = iend;
    riord[iend++] = ii;
    mask[ii] = 0;
  }
  levels[nlev] = iend;
  for(int j = 0; j < iend; j++) mask[riord[j]] = 1;
  return nlev;
}

int orc_rcmk(int n, const int *ia, const int *ja, const double *a, int *perm,
             int *rperm, int *ia_out, int *ja_out, double *a_out)
{
  int *mask = (int *)malloc(sizeof(int) * (size_t)n);
  int *levels = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  for(int i = 0; i < n; i++) mask[i] = 1;
  /* perphn, reordering.f:232-271 (init = node 1) */
  int init = 0, nlevp = 0, nlev;
  for(;;) {
    nlev = bfs_levels(n, ia, ja, init, mask, rperm, levels);
    if(nlev <= nlevp) break;
    int mindeg = levels[nlev];
    for(int j = levels[nlev - 1]; j < levels[nlev]; j++) {
      int nod = rperm[j], deg = 0;
      for(int k = ia[nod]; k < ia[nod + 1]; k++)
        if(mask[ja[k]] == 1) deg++; /* maskdeg, reordering.f:379-391 */
      if(deg < mindeg) {
        init = nod;
        mindeg = deg;
      }
    }
    nlevp = nlev;
  }
  /* rversp + exchange */
  for(int j = 0; j < n / 2; j++) {
    int k = rperm[j];
    rperm[j] = rperm[n - 1 - j];
    rperm[n - 1 - j] = k;
  }
  for(int i = 0; i < n; i++) perm[rperm[i]] = i;
  /* dperm (rperm then cperm with the same permutation) */
  ia_out[0] = 0;
  for(int j = 0; j < n; j++) ia_out[perm[j] + 1] = ia[j + 1] - ia[j];
  for(int j = 0; j < n; j++) ia_out[j + 1] += ia_out[j];
  for(int ii = 0; ii < n; ii++) {
    int ko = ia_out[perm[ii]];
    for(int k = ia[ii]; k < ia[ii + 1]; k++, ko++) {
      ja_out[ko] = perm[ja[k]];
      a_out[ko] = a[k];
    }
  }
  /* sortcol (cmkreord.f:88-112): per-row selection sort by column */
  for(int i = 0; i < n; i++) {
    for(int p = ia_out[i]; p < ia_out[i + 1]; p++) {
      int jmin = p;
      for(int q = p + 1; q < ia_out[i + 1]; q++)
        if(ja_out[q] < ja_out[jmin]) jmin = q;
      int tj = ja_out[p];
      ja_out[p] = ja_out[jmin];
      ja_out[jmin] = tj;
      double ta = a_out[p];
      a_out[p] = a_out[jmin];
      a_out[jmin] = ta;
    }
  }
  free(mask);
  free(levels);
  return nlev;
}

/* ---- solve_matrix (Sparskit.cpp:108-634) -------------------------------- */

typedef struct {
  int n;
  const int *ia, *ja;
  const double *a;
  int prec; /* 0 none, 6 ILU0, 8 DIAGONAL */
  int f64;
  const float *alu32;
  const double *alu64;
  const int *jlu, *ju;
} solve_ctx;

static void op_amux(void *c, const double *in, double *out)
{
  solve_ctx *s = (solve_ctx *)c;
  orc_amux(s->n, in, out, s->ia, s->ja, s->a);
}

static void op_lusol(void *c, const double *in, double *out)
{
  solve_ctx *s = (solve_ctx *)c;
  if(s->f64) orc_lusol_f64(s->n, in, out, s->alu64, s->jlu, s->ju);
  else orc_lusol_f32(s->n, in, out, s->alu32, s->jlu, s->ju);
}

int orc_solve(int n, const int *ia, const int *ja, double *a, double *b,
              double *x, const orc_params *p, orc_info *info)
{
  int nnz = ia[n];
  double *rowscal = NULL;
  memset(info, 0, sizeof(*info));
  if(p->scaling == 1) { /* scale_matrix DIAG_SCALING, Sparskit.cpp:975-995 */
    rowscal = (double *)malloc(sizeof(double) * (size_t)n);
    for(int i = 0; i < n; i++) {
      rowscal[i] = 0.0;
      for(int k = ia[i]; k < ia[i + 1]; k++)
        if(ja[k] == i) rowscal[i] = a[k];
    }
    for(int i = 0; i < n; i++) {
      if(rowscal[i]) rowscal[i] = 1. / sqrt(fabs(rowscal[i]));
      else {
        free(rowscal);
        rowscal = NULL;
        break;
      }
    }
    if(rowscal) {
      for(int i = 0; i < n; i++)
        for(int k = ia[i]; k < ia[i + 1]; k++) a[k] = a[k] * rowscal[i];
      for(int i = 0; i < n; i++)
        for(int k = ia[i]; k < ia[i + 1]; k++) a[k] = a[k] * rowscal[ja[k]];
      for(int i = 0; i < n; i++) b[i] *= rowscal[i];
    }
  }
  for(int i = 1; i < n; i++)
    if(ia[i] - ia[i - 1] <= 0) {
      free(rowscal);
      return -100;
    }

  int *perm = (int *)malloc(sizeof(int) * (size_t)n);
  int *rperm = (int *)malloc(sizeof(int) * (size_t)n);
  const int *IA = ia, *JA = ja;
  const double *A = a;
  int *ia_r = NULL, *ja_r = NULL;
  double *a_r = NULL;
  for(int i = 0; i < n; i++) perm[i] = rperm[i] = i;
  if(p->renumbering == 1) {
    ia_r = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    ja_r = (int *)malloc(sizeof(int) * (size_t)nnz);
    a_r = (double *)malloc(sizeof(double) * (size_t)nnz);
    orc_rcmk(n, ia, ja, a, perm, rperm, ia_r, ja_r, a_r);
    IA = ia_r;
    JA = ja_r;
    A = a_r;
  }

  solve_ctx sc;
  memset(&sc, 0, sizeof(sc));
  sc.n = n;
  sc.ia = IA;
  sc.ja = JA;
  sc.a = A;
  sc.prec = p->preconditioner;
  sc.f64 = p->ilu_double;
  float *alu32 = NULL;
  double *alu64 = NULL;
  int *jlu = NULL, *ju = NULL;
  int ret = 0;
  int *ja_p = NULL, *iperm = NULL; /* ILUTP: permuted columns, permutation */
  if(p->preconditioner == 1 || p->preconditioner == 2) {
    /* Sparskit.cpp:312-317, 375-390, 435-446: first guess of the factor size,
     * grown by half and factored again on -2/-3; ilutp permutes the columns of
     * the matrix the iteration then multiplies with */
    int nnz_ilu = 2 * (n + 1) * (p->nb_fill + 1);
    ja_p = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    iperm = (int *)malloc(sizeof(int) * 2 * (size_t)(n > 0 ? n : 1));
    for(;;) {
      memcpy(ja_p, JA, sizeof(int) * (size_t)nnz);
      jlu = (int *)malloc(sizeof(int) * ((size_t)nnz_ilu + 2));
      ju = (int *)malloc(sizeof(int) * (size_t)(n + 1));
      int ierr;
      if(p->ilu_double) {
        alu64 = (double *)malloc(sizeof(double) * ((size_t)nnz_ilu + 2));
        ierr = orc_ilut_f64(n, IA, ja_p, A, p->nb_fill, p->droptol, p->permtol,
                            p->preconditioner == 2, nnz_ilu, alu64, jlu, ju, iperm);
      }
      else {
        alu32 = (float *)malloc(sizeof(float) * ((size_t)nnz_ilu + 2));
        ierr = orc_ilut_f32(n, IA, ja_p, A, p->nb_fill, p->droptol, p->permtol,
                            p->preconditioner == 2, nnz_ilu, alu32, jlu, ju, iperm);
      }
      if(ierr == -2 || ierr == -3) {
        nnz_ilu += nnz_ilu / 2;
        free(alu32);
        free(alu64);
        free(jlu);
        free(ju);
        alu32 = NULL;
        alu64 = NULL;
        continue;
      }
      info->ilu_ierr = ierr;
      break;
    }
    if(p->preconditioner == 1)
      for(int i = 0; i < n; i++) iperm[n + i] = i;
    JA = ja_p;
    sc.ja = JA;
    sc.alu32 = alu32;
    sc.alu64 = alu64;
    sc.jlu = jlu;
    sc.ju = ju;
    ret = info->ilu_ierr;
  }
  else if(p->preconditioner == 6) {
    size_t nz = (size_t)nnz + n + 1;
    jlu = (int *)malloc(sizeof(int) * nz);
    ju = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    if(p->ilu_double) {
      alu64 = (double *)malloc(sizeof(double) * nz);
      info->ilu_ierr = orc_ilu0_f64(n, IA, JA, A, alu64, jlu, ju);
    }
    else {
      alu32 = (float *)malloc(sizeof(float) * nz);
      info->ilu_ierr = orc_ilu0_f32(n, IA, JA, A, alu32, jlu, ju);
    }
    sc.alu32 = alu32;
    sc.alu64 = alu64;
    sc.jlu = jlu;
    sc.ju = ju;
    ret = info->ilu_ierr;
  }
  if(!ret) {
    double *rhs = (double *)malloc(sizeof(double) * (size_t)n);
    double *sol = (double *)calloc((size_t)n, sizeof(double));
    for(int i = 0; i < n; i++) rhs[i] = b[rperm[i]];
    /* DIAGONAL (=8) is the identity (Sparskit.cpp:338-354): applying it is a
     * copy, which leaves every value unchanged. */
    orc_op prec = NULL;
    if(p->preconditioner == 6 || p->preconditioner == 1 || p->preconditioner == 2) prec = op_lusol;
    int pos = p->precond_position ? p->precond_position : 2;
    if(pos != 2 && prec) prec = NULL; /* left preconditioning: not restated */
    static const int dummy = 0;
    (void)dummy;
    if(p->preconditioner == 8) prec = NULL;
    switch(p->algorithm) {
    case 1:
      orc_cg(n, rhs, sol, p->maxits, p->rtol, op_amux, prec, &sc, info);
      break;
    case 5:
      orc_bcgstab(n, rhs, sol, p->maxits, p->rtol, op_amux, prec, &sc, info);
      break;
    case 6:
      orc_tfqmr(n, rhs, sol, p->maxits, p->rtol, op_amux, prec, &sc, info);
      break;
    case 9:
      orc_fgmres(n, rhs, sol, p->krylov, p->maxits, p->rtol, op_amux, prec, &sc,
                 info);
      break;
    default:
      orc_gmres(n, rhs, sol, p->krylov, p->maxits, p->rtol, op_amux, prec, &sc,
                info);
      break;
    }
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    orc_amux(n, sol, w, IA, JA, A);
    double s2 = 0.0;
    for(int i = 0; i < n; i++) {
      w[i] -= rhs[i];
      s2 += w[i] * w[i];
    }
    info->true_res = sqrt(s2);
    for(int i = 0; i < n; i++) /* Sparskit.cpp:597-602 */
      x[i] = iperm ? sol[iperm[n + perm[i]]] : sol[perm[i]];
    if(rowscal)
      for(int i = 0; i < n; i++) x[i] *= rowscal[i];
    free(w);
    free(rhs);
    free(sol);
  }
  free(perm);
  free(rperm);
  free(ia_r);
  free(ja_r);
  free(a_r);
  free(alu32);
  free(alu64);
  free(jlu);
  free(ju);
  free(ja_p);
  free(iperm);
  free(rowscal);
  return ret;
}

/* ========================================================================
 * Mesh entities
 * ====================================================================== */

/* Local edge tables Den_Triangle / Den_Tetrahedron (GeoEntity.h:93-94,
 * 214-215): local edge k joins local nodes (ea[k], eb[k]). */
static const int TRI_EA[3] = {0, 0, 1}, TRI_EB[3] = {1, 2, 2};
static const int TET_EA[6] = {0, 0, 0, 1, 1, 2}, TET_EB[6] = {1, 2, 3, 2, 3, 3};

static int nedges_of(int t) { return t == ORC_TET ? 6 : t == ORC_TRI ? 3 : t == ORC_LINE ? 1 : 0; }
static int nnodes_of(int t) { return t == ORC_TET ? 4 : t == ORC_TRI ? 3 : t == ORC_LINE ? 2 : 0; }

typedef struct {
  long long key;
  int num;
} edge_slot;

static unsigned long long hash64(unsigned long long k)
{
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

int orc_number_edges(int ne, const int *etype, const int *conn, int nvisit,
                     const int *visit, int *edges_out, int *edge_nodes,
                     int max_edges)
{
  size_t cap = 16;
  while(cap < (size_t)max_edges * 2 + 16) cap <<= 1;
  edge_slot *tab = (edge_slot *)calloc(cap, sizeof(edge_slot));
  int count = 0;
  memset(edges_out, 0, sizeof(int) * (size_t)ne * 6);
  for(int vi = 0; vi < nvisit; vi++) {
    int e = visit[vi];
    int t = etype[e];
    const int *ea = t == ORC_TET ? TET_EA : TRI_EA;
    const int *eb = t == ORC_TET ? TET_EB : TRI_EB;
    int ned = nedges_of(t);
    if(t == ORC_LINE) {
      static const int LA[1] = {0}, LB[1] = {1};
      ea = LA;
      eb = LB;
    }
    if(edges_out[(size_t)e * 6] != 0) continue; /* already created */
    for(int k = 0; k < ned; k++) {
      int na = conn[(size_t)e * 4 + ea[k]], nb = conn[(size_t)e * 4 + eb[k]];
      int lo = na < nb ? na : nb, hi = na < nb ? nb : na;
      int sign = na < nb ? 1 : -1;
      long long key = ((long long)lo << 32) | (unsigned)hi;
      size_t h = hash64((unsigned long long)key) & (cap - 1);
      while(tab[h].num && tab[h].key != key) h = (h + 1) & (cap - 1);
      if(!tab[h].num) {
        if(count >= max_edges) {
          free(tab);
          return -1;
        }
        tab[h].key = key;
        tab[h].num = ++count;
        if(edge_nodes) {
          edge_nodes[2 * (count - 1)] = lo;
          edge_nodes[2 * (count - 1) + 1] = hi;
        }
      }
      edges_out[(size_t)e * 6 + k] = sign * tab[h].num;
    }
  }
  free(tab);
  return count;
}

/* ========================================================================
 * Element math
 * ====================================================================== */

typedef struct {
  double c11, c12, c13, c21, c22, c23, c31, c32, c33;
} m3;

/* Gauss tables with the reference's truncated literals
 * (Gauss_Triangle.h:8-22, Gauss_Tetrahedron.h:8-23). */
static int gauss_point(int t, int np, int q, double *u, double *v, double *w,
                       double *wt)
{
  static const double xt3[3] = {0.16666666666666, 0.66666666666666, 0.16666666666666};
  static const double yt3[3] = {0.16666666666666, 0.16666666666666, 0.66666666666666};
  static const double pt3[3] = {0.16666666666666, 0.16666666666666, 0.16666666666666};
  static const double xt4[4] = {0.333333333333333, 0.6, 0.2, 0.2};
  static const double yt4[4] = {0.3333333333333333, 0.2, 0.6, 0.2};
  static const double pt4[4] = {-0.28125, .260416666666, .260416666666, .260416666666};
  static const double xtet4[4] = {0.138196601125, 0.138196601125, 0.138196601125, 0.585410196625};
  static const double ytet4[4] = {0.138196601125, 0.138196601125, 0.585410196625, 0.138196601125};
  static const double ztet4[4] = {0.138196601125, 0.585410196625, 0.138196601125, 0.138196601125};
  static const double ptet4[4] = {0.0416666666667, 0.0416666666667, 0.0416666666667, 0.0416666666667};
  *w = 0.;
  if(t == ORC_TRI) {
    if(np == 1) { *u = 0.333333333333333; *v = 0.333333333333333; *wt = 0.5; return 0; }
    if(np == 3) { *u = xt3[q]; *v = yt3[q]; *wt = pt3[q]; return 0; }
    if(np == 4) { *u = xt4[q]; *v = yt4[q]; *wt = pt4[q]; return 0; }
  }
  else if(t == ORC_TET) {
    if(np == 1) { *u = *v = *w = 0.25; *wt = 0.166666666667; return 0; }
    if(np == 4) { *u = xtet4[q]; *v = ytet4[q]; *w = ztet4[q]; *wt = ptet4[q]; return 0; }
  }
  return -1;
}

/* BF_Node / BF_GradNode, BF_Node.cpp:49-76, 245-320. */
static double bf_node(int t, int i, double u, double v, double w)
{
  if(t == ORC_TRI) return i == 0 ? 1. - u - v : i == 1 ? u : v;
  return i == 0 ? 1. - u - v - w : i == 1 ? u : i == 2 ? v : w;
}

static void bf_gradnode(int t, int i, double s[3])
{
  s[0] = s[1] = s[2] = 0.;
  if(i == 0) {
    s[0] = -1.;
    s[1] = -1.;
    s[2] = (t == ORC_TET) ? -1. : 0.;
  }
  else
    s[i - 1] = 1.;
}

/* BF_Edge / BF_CurlEdge for tetrahedra in reference coordinates,
 * BF_Edge.cpp:95-131 and :452-486 (unsigned; sign applied by the caller as at
 * BF_Edge.cpp:366-372, 720-726). */
static void bf_edge_tet(int k, double u, double v, double w, double s[3])
{
  switch(k) {
  case 0: s[0] = 1. - v - w; s[1] = u; s[2] = u; break;
  case 1: s[0] = v; s[1] = 1. - u - w; s[2] = v; break;
  case 2: s[0] = w; s[1] = w; s[2] = 1. - u - v; break;
  case 3: s[0] = -v; s[1] = u; s[2] = 0.; break;
  case 4: s[0] = -w; s[1] = 0.; s[2] = u; break;
  default: s[0] = 0.; s[1] = -w; s[2] = v; break;
  }
}

static void bf_curledge_tet(int k, double s[3])
{
  static const double C[6][3] = {{0., -2., 2.}, {2., 0., -2.}, {-2., 2., 0.},
                                 {0., 0., 2.},  {0., -2., 0.}, {2., 0., 0.}};
  s[0] = C[k][0];
  s[1] = C[k][1];
  s[2] = C[k][2];
}

/* BF_CurlPerpendicularEdge (BF_Perpendicular.cpp:70-78): the node gradient
 * in reference coordinates turned by -90 degrees, (dv, -du, dw). */
static void bf_curlperp(int t, int i, double s[3])
{
  bf_gradnode(t, i, s);
  double ss = s[0];
  s[0] = s[1];
  s[1] = -ss;
}

/* JacobianVol2D / JacobianVol3D / JacobianSur3D, Get_Geometry.cpp:1099-1124,
 * 1435-1469, 1629-1680.  dndu are the P1 gradients (constant). */
static double jacobian(int t, int sur, const double *x, const double *y,
                       const double *z, m3 *J)
{
  int nn = nnodes_of(t);
  double g[4][3];
  for(int i = 0; i < nn; i++) bf_gradnode(t, i, g[i]);
  memset(J, 0, sizeof(*J));
  if(t == ORC_TRI && !sur) {
    J->c33 = 1.;
    for(int i = 0; i < nn; i++) {
      J->c11 += x[i] * g[i][0];
      J->c21 += x[i] * g[i][1];
      J->c12 += y[i] * g[i][0];
      J->c22 += y[i] * g[i][1];
    }
    return J->c11 * J->c22 - J->c12 * J->c21;
  }
  if(t == ORC_TRI && sur) {
    for(int i = 0; i < nn; i++) {
      J->c11 += x[i] * g[i][0];
      J->c21 += x[i] * g[i][1];
      J->c12 += y[i] * g[i][0];
      J->c22 += y[i] * g[i][1];
      J->c13 += z[i] * g[i][0];
      J->c23 += z[i] * g[i][1];
    }
    double a1 = J->c11 * J->c22 - J->c12 * J->c21;
    double a2 = J->c13 * J->c21 - J->c11 * J->c23;
    double a3 = J->c12 * J->c23 - J->c13 * J->c22;
    return sqrt(a1 * a1 + a2 * a2 + a3 * a3);
  }
  for(int i = 0; i < nn; i++) {
    J->c11 += x[i] * g[i][0];
    J->c21 += x[i] * g[i][1];
    J->c31 += x[i] * g[i][2];
    J->c12 += y[i] * g[i][0];
    J->c22 += y[i] * g[i][1];
    J->c32 += y[i] * g[i][2];
    J->c13 += z[i] * g[i][0];
    J->c23 += z[i] * g[i][1];
    J->c33 += z[i] * g[i][2];
  }
  return J->c11 * J->c22 * J->c33 + J->c13 * J->c21 * J->c32 +
         J->c12 * J->c23 * J->c31 - J->c13 * J->c22 * J->c31 -
         J->c11 * J->c23 * J->c32 - J->c12 * J->c21 * J->c33;
}

/* Get_InverseMatrix, Get_Geometry.cpp:1755-1808. */
static void inverse(int dim3, double d, const m3 *M, m3 *I)
{
  if(!dim3) {
    I->c11 = M->c22 * M->c33 / d;
    I->c21 = -M->c21 * M->c33 / d;
    I->c12 = -M->c12 * M->c33 / d;
    I->c22 = M->c11 * M->c33 / d;
    I->c13 = I->c23 = I->c31 = I->c32 = 0.;
    I->c33 = 1. / M->c33;
    return;
  }
  I->c11 = (M->c22 * M->c33 - M->c23 * M->c32) / d;
  I->c21 = -(M->c21 * M->c33 - M->c23 * M->c31) / d;
  I->c31 = (M->c21 * M->c32 - M->c22 * M->c31) / d;
  I->c12 = -(M->c12 * M->c33 - M->c13 * M->c32) / d;
  I->c22 = (M->c11 * M->c33 - M->c13 * M->c31) / d;
  I->c32 = -(M->c11 * M->c32 - M->c12 * M->c31) / d;
  I->c13 = (M->c12 * M->c23 - M->c13 * M->c22) / d;
  I->c23 = -(M->c11 * M->c23 - M->c13 * M->c21) / d;
  I->c33 = (M->c11 * M->c22 - M->c12 * M->c21) / d;
}

/* ChangeOfCoord_Form1 / Form2, Get_Geometry.cpp:1895-1925. */
static void form1(const m3 *I, const double u[3], double x[3])
{
  x[0] = u[0] * I->c11 + u[1] * I->c12 + u[2] * I->c13;
  x[1] = u[0] * I->c21 + u[1] * I->c22 + u[2] * I->c23;
  x[2] = u[0] * I->c31 + u[1] * I->c32 + u[2] * I->c33;
}

static void form2(const m3 *J, double d, const double u[3], double x[3])
{
  x[0] = (u[0] * J->c11 + u[1] * J->c21 + u[2] * J->c31) / d;
  x[1] = (u[0] * J->c12 + u[1] * J->c22 + u[2] * J->c32) / d;
  x[2] = (u[0] * J->c13 + u[1] * J->c23 + u[2] * J->c33) / d;
}

/* ChangeOfCoord_Form1P / Form2P (Get_Geometry.cpp:1931-1947), 2D only. */
static void form1p(const m3 *J, const double u[3], double x[3])
{
  x[0] = 0.;
  x[1] = 0.;
  x[2] = u[2] / J->c33;
}

static void form2p(const m3 *J, double d, const double u[3], double x[3])
{
  x[0] = (u[0] * J->c11 + u[1] * J->c21) / d;
  x[1] = (u[0] * J->c12 + u[1] * J->c22) / d;
  x[2] = 0.;
}

/* Cal_ProductValue (Cal_Value.cpp:626-900): coefficient times the DOF-side
 * value; scalar x scalar, scalar x vector, tensor x vector. */
static void coef_product(int ncoef, const double *c, const double v[3],
                         double r[3])
{
  if(ncoef == 1) {
    double a0 = c[0];
    r[0] = a0 * v[0];
    r[1] = a0 * v[1];
    r[2] = a0 * v[2];
  }
  else {
    double t0 = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
    double t1 = c[3] * v[0] + c[4] * v[1] + c[5] * v[2];
    double t2 = c[6] * v[0] + c[7] * v[1] + c[8] * v[2];
    r[0] = t0;
    r[1] = t1;
    r[2] = t2;
  }
}

#define MAXK 12

static int term_is_symmetric(const orc_term *t)
{
  /* Cal_GalerkinTermOfFemEquation.cpp:72-99: same quantity and operator */
  return t->kind < 10 && t->equ == t->dof;
}

static int term_nbf(const orc_term *t, int perp, int etype)
{
  if(!perp && (t->kind == ORC_T_CURL_CURL || t->kind == ORC_T_EDGE_MASS ||
               t->kind == ORC_T_SRC_EDGE))
    return nedges_of(etype);
  return nnodes_of(etype);
}

/* Element matrix of one term.  perp: the test space is BF_PerpendicularEdge
 * (triangles in the plane only). */
static int elem_matrix_impl(const orc_mesh *m, const orc_term *t, int perp, int e,
                            double Ek[MAXK][MAXK], int *kequ, int *kdof)
{
  int et = m->etype[e];
  int nn = nnodes_of(et);
  double x[4], y[4], z[4];
  for(int i = 0; i < nn; i++) {
    int nd = m->conn[(size_t)e * 4 + i] - 1;
    x[i] = m->xyz[3 * (size_t)nd];
    y[i] = m->xyz[3 * (size_t)nd + 1];
    z[i] = m->xyz[3 * (size_t)nd + 2];
  }
  int reg = m->region[e];
  if(reg < 0 || reg >= t->coef_nreg) return -1;
  const double *coef = t->coef + (size_t)reg * t->ncoef;
  int sym = term_is_symmetric(t);
  if(perp && (et != ORC_TRI || t->jac_sur)) return -4;
  int ke = term_nbf(t, perp, et);
  int kd = t->kind >= 10 ? 1 : ke;
  int is_src = t->kind >= 10;
  int dim3 = (et == ORC_TET) || t->jac_sur;
  *kequ = ke;
  *kdof = kd;
  for(int i = 0; i < MAXK; i++)
    for(int j = 0; j < MAXK; j++) Ek[i][j] = 0.;
  int esign[6] = {1, 1, 1, 1, 1, 1};
  if(!perp && (t->kind == ORC_T_CURL_CURL || t->kind == ORC_T_EDGE_MASS ||
                t->kind == ORC_T_SRC_EDGE)) {
    if(!m->edges) return -2;
    for(int k = 0; k < ke; k++) esign[k] = m->edges[(size_t)e * 6 + k] < 0 ? -1 : 1;
  }
  for(int q = 0; q < t->npoints; q++) {
    double u, v, w, wt;
    if(gauss_point(et, t->npoints, q, &u, &v, &w, &wt)) return -3;
    m3 J, I;
    double det = jacobian(et, t->jac_sur, x, y, z, &J);
    int need_inv = !perp && (t->kind == ORC_T_GRAD_GRAD || t->kind == ORC_T_EDGE_MASS ||
                             t->kind == ORC_T_SRC_EDGE);
    if(need_inv) inverse(dim3, det, &J, &I);
    double be[MAXK][3], bd[MAXK][3];
    for(int i = 0; i < ke; i++) {
      double s[3];
      if(perp) {
        if(t->kind == ORC_T_CURL_CURL) {
          bf_curlperp(et, i, s);
          form2p(&J, det, s, be[i]);
        }
        else { /* BF_PerpendicularEdge value (0, 0, BF_Node), :33-41 */
          s[0] = s[1] = 0.;
          s[2] = bf_node(et, i, u, v, w);
          form1p(&J, s, be[i]);
        }
        continue;
      }
      switch(t->kind) {
      case ORC_T_GRAD_GRAD:
        bf_gradnode(et, i, s);
        form1(&I, s, be[i]);
        break;
      case ORC_T_CURL_CURL:
        bf_curledge_tet(i, s);
        for(int c = 0; c < 3; c++) s[c] *= esign[i];
        form2(&J, det, s, be[i]);
        break;
      case ORC_T_EDGE_MASS:
      case ORC_T_SRC_EDGE:
        bf_edge_tet(i, u, v, w, s);
        for(int c = 0; c < 3; c++) s[c] = esign[i] < 0 ? -s[c] : s[c];
        form1(&I, s, be[i]);
        break;
      default: /* node value (FORM0): ChangeOfCoord_No1 */
        be[i][0] = bf_node(et, i, u, v, w);
        be[i][1] = be[i][2] = 0.;
        break;
      }
    }
    if(is_src) {
      bd[0][0] = coef[0];
      bd[0][1] = t->ncoef > 1 ? coef[1] : 0.;
      bd[0][2] = t->ncoef > 2 ? coef[2] : 0.;
    }
    else {
      /* symmetric: the DOF side is the transformed test side (no change of
       * coordinates, Cal_GalerkinTermOfFemEquation.cpp:484-492); otherwise
       * it is evaluated anew -- the values are the same for these forms. */
      for(int j = 0; j < kd; j++) {
        double tmp[3] = {be[j][0], be[j][1], be[j][2]};
        if(t->kind == ORC_T_NODE_MASS) {
          bd[j][0] = coef[0] * tmp[0];
          bd[j][1] = bd[j][2] = 0.;
        }
        else
          coef_product(t->ncoef, coef, tmp, bd[j]);
      }
    }
    double factor = 1. * (wt * fabs(det));
    int prod1 = (t->kind == ORC_T_NODE_MASS || t->kind == ORC_T_SRC_NODE);
    /* Cal_Product3 for FORM1P test functions */
    int prod3 = perp && (t->kind == ORC_T_EDGE_MASS || t->kind == ORC_T_SRC_EDGE);
    for(int i = 0; i < ke; i++) {
      int jmax = sym ? i + 1 : kd;
      for(int j = 0; j < jmax; j++) {
        double pr = prod1 ? be[i][0] * bd[j][0]
                    : prod3 ? be[i][2] * bd[j][2]
                            : be[i][0] * bd[j][0] + be[i][1] * bd[j][1] +
                                be[i][2] * bd[j][2];
        Ek[i][j] += factor * pr;
      }
    }
  }
  if(sym)
    for(int i = 1; i < ke; i++)
      for(int j = 0; j < i; j++) Ek[j][i] = Ek[i][j];
  return 0;
}

int orc_element_matrix(const orc_mesh *m, const orc_space *sp, const orc_term *t, int e,
                       double *Ek, int *k_out)
{
  double E[MAXK][MAXK];
  int ke, kd;
  int r = elem_matrix_impl(m, t, sp[t->equ].bf == ORC_BF_PERP_EDGE, e, E, &ke, &kd);
  if(r) return r;
  for(int i = 0; i < ke; i++)
    for(int j = 0; j < kd; j++) Ek[i * kd + j] = E[i][j];
  k_out[0] = ke;
  k_out[1] = kd;
  return 0;
}

/* ========================================================================
 * DOF numbering and assembly
 * ====================================================================== */

static int in_list(int v, int n, const int *l)
{
  for(int i = 0; i < n; i++)
    if(l[i] == v) return 1;
  return 0;
}

/* A term contributes on element e when e's region is among the term's regions
 * and in the supports of the spaces it uses (Get_DofOfElement.cpp:102-110
 * gathers no basis function outside the support). */
static int term_active(const orc_mesh *m, const orc_space *sp, const orc_term *t, int e)
{
  int r = m->region[e];
  if(!in_list(r, t->nregions, t->regions)) return 0;
  if(!in_list(r, sp[t->equ].nsupport, sp[t->equ].support)) return 0;
  if(t->dof >= 0 && !in_list(r, sp[t->dof].nsupport, sp[t->dof].support)) return 0;
  return 1;
}

static int space_entity(const orc_mesh *m, const orc_space *s, int e, int k,
                        int *sign)
{
  *sign = 1;
  if(s->bf != ORC_BF_EDGE) return m->conn[(size_t)e * 4 + k];
  int v = m->edges[(size_t)e * 6 + k];
  *sign = v < 0 ? -1 : 1;
  return v < 0 ? -v : v;
}

int orc_number_dofs(const orc_mesh *m, int nspaces, const orc_space *sp,
                    int nterms, const orc_term *t, int nmax, int *dofnum)
{
  memset(dofnum, 0, sizeof(int) * (size_t)nspaces * (size_t)(nmax + 1));
  /* mark (space, entity) pairs touched by the terms (PRE pass,
   * Pre_TermOfFemEquation.cpp:125-160) */
  for(int e = 0; e < m->nelem; e++) {
    for(int ti = 0; ti < nterms; ti++) {
      const orc_term *tt = &t[ti];
      if(!term_active(m, sp, tt, e)) continue;
      int sps[2] = {tt->equ, tt->dof};
      for(int q = 0; q < 2; q++) {
        int s = sps[q];
        if(s < 0) continue;
        int nb = sp[s].bf != ORC_BF_EDGE ? nnodes_of(m->etype[e]) : nedges_of(m->etype[e]);
        for(int k = 0; k < nb; k++) {
          int sg;
          int ent = space_entity(m, &sp[s], e, k, &sg);
          if(ent < 1 || ent > nmax) return -1;
          dofnum[(size_t)s * (nmax + 1) + ent] = 1;
        }
      }
    }
  }
  /* fixed DOFs: the first constraint that holds the entity wins
   * (Treatment_ConstraintForElement stops at the first match,
   * Get_ConstraintOfElement.cpp:51-54), and Dof_DefineAssignFixedDof only
   * turns an unknown DOF into a fixed one, never re-assigns a fixed one
   * (DofData.cpp:1137-1151) */
  for(int s = 0; s < nspaces; s++)
    for(int f = 0; f < sp[s].nfixed; f++) {
      int ent = sp[s].fixed_entity[f];
      if(ent >= 1 && ent <= nmax && dofnum[(size_t)s * (nmax + 1) + ent] > 0)
        dofnum[(size_t)s * (nmax + 1) + ent] = -(f + 1);
    }
  /* number unknowns in (NumType, Entity) order: AVL in-order traversal,
   * DofData.cpp:1299-1323 with fcmp_Dof :36-45 */
  int order[64];
  for(int s = 0; s < nspaces; s++) order[s] = s;
  for(int a = 0; a < nspaces; a++)
    for(int b = a + 1; b < nspaces; b++)
      if(sp[order[b]].numtype < sp[order[a]].numtype) {
        int tmp = order[a];
        order[a] = order[b];
        order[b] = tmp;
      }
  int N = 0;
  for(int oi = 0; oi < nspaces; oi++) {
    int s = order[oi];
    for(int ent = 1; ent <= nmax; ent++) {
      int *d = &dofnum[(size_t)s * (nmax + 1) + ent];
      if(*d == 1) *d = ++N;
      else if(*d > 0) *d = ++N;
    }
  }
  return N;
}

typedef struct {
  int n, cap;
  int *col;
  double *val;
} row_list;

static void row_add(row_list *r, int col, double v)
{
  for(int k = 0; k < r->n; k++)
    if(r->col[k] == col) {
      r->val[k] += v;
      return;
    }
  if(r->n == r->cap) {
    r->cap = r->cap ? 2 * r->cap : 8;
    r->col = (int *)realloc(r->col, sizeof(int) * (size_t)r->cap);
    r->val = (double *)realloc(r->val, sizeof(double) * (size_t)r->cap);
  }
  r->col[r->n] = col;
  r->val[r->n] = v;
  r->n++;
}

/* CSR of the row lists (csr_format, Sparskit.cpp:2186-2208): rows by
 * equation, columns sorted; frees the lists.  Returns nnz. */
static long rows_to_csr(row_list *rows, int N, int *ia, int *ja, double *a, long max_nnz, int threads)
{
  #pragma omp parallel for schedule(dynamic, 1024) if(threads > 1) num_threads(threads)
  for(int i = 0; i < N; i++) {
    row_list *r = &rows[i];
    for(int p = 1; p < r->n; p++) { /* insertion sort by column */
      int c = r->col[p];
      double v = r->val[p];
      int q = p - 1;
      while(q >= 0 && r->col[q] > c) {
        r->col[q + 1] = r->col[q];
        r->val[q + 1] = r->val[q];
        q--;
      }
      r->col[q + 1] = c;
      r->val[q + 1] = v;
    }
  }
  long nnz = 0;
  ia[0] = 0;
  for(int i = 0; i < N; i++) {
    nnz += rows[i].n;
    ia[i + 1] = (int)nnz;
  }
  #pragma omp parallel for schedule(dynamic, 1024) if(threads > 1) num_threads(threads)
  for(int i = 0; i < N; i++) {
    row_list *r = &rows[i];
    for(int p = 0; p < r->n; p++) {
      long k = (long)ia[i] + p;
      if(k < max_nnz) {
        ja[k] = r->col[p];
        a[k] = r->val[p];
      }
    }
    free(r->col);
    free(r->val);
  }
  free(rows);
  return nnz;
}

/* The element loop on g_threads threads, same result as the sequential one
 * bit for bit: the element matrices of a block of elements are formed in
 * parallel, then each thread adds, in element order, the contributions to
 * its own contiguous range of equations (matrix rows and right-hand side). */
static long assemble_threaded(const orc_mesh *m, const orc_space *sp, int nterms, const orc_term *t,
                              int nmax, const int *dofnum, int N, int *ia, int *ja, double *a,
                              long max_nnz, double *b)
{
  const int T = g_threads;
  const long blk = 32768;
  const size_t per = (size_t)MAXK * MAXK;
  const long npair = blk * nterms;
  double *Ek = (double *)malloc(sizeof(double) * per * (size_t)npair);
  int *kk = (int *)malloc(sizeof(int) * 2 * (size_t)npair);
  row_list *rows = (row_list *)calloc((size_t)N, sizeof(row_list));
  int err = 0;
  for(int i = 0; i < N; i++) b[i] = 0.;
  for(long e0 = 0; e0 < m->nelem && !err; e0 += blk) {
    const long ne = m->nelem - e0 < blk ? m->nelem - e0 : blk;
    #pragma omp parallel for schedule(dynamic, 256) num_threads(T)
    for(long q = 0; q < ne * nterms; q++) {
      const int e = (int)(e0 + q / nterms);
      const orc_term *tt = &t[q % nterms];
      kk[2 * q] = -1;
      if(!term_active(m, sp, tt, e)) continue;
      if(elem_matrix_impl(m, tt, sp[tt->equ].bf == ORC_BF_PERP_EDGE, e, (double(*)[MAXK])(Ek + per * (size_t)q),
                          &kk[2 * q], &kk[2 * q + 1])) {
        #pragma omp atomic write
        err = 1;
      }
    }
    if(err) break;
    #pragma omp parallel num_threads(T)
    {
      const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
      const int r0 = (int)((long)N * tid / nt), r1 = (int)((long)N * (tid + 1) / nt);
      for(long q = 0; q < ne * nterms; q++) {
        const int ke = kk[2 * q], kd = kk[2 * q + 1];
        if(ke < 0) continue;
        const int e = (int)(e0 + q / nterms);
        const orc_term *tt = &t[q % nterms];
        const double(*E)[MAXK] = (const double(*)[MAXK])(Ek + per * (size_t)q);
        for(int i = 0; i < ke; i++) {
          int sg;
          int eq_ent = space_entity(m, &sp[tt->equ], e, i, &sg);
          int eq = dofnum[(size_t)tt->equ * (nmax + 1) + eq_ent];
          if(eq <= 0 || eq - 1 < r0 || eq - 1 >= r1) continue;
          for(int j = 0; j < kd; j++) {
            double val = E[i][j];
            if(tt->dof < 0) {
              if(val) b[eq - 1] += (1. * 1.) * (-val);
              continue;
            }
            int dent = space_entity(m, &sp[tt->dof], e, j, &sg);
            int dn = dofnum[(size_t)tt->dof * (nmax + 1) + dent];
            if(dn > 0) row_add(&rows[eq - 1], dn - 1, val);
            else if(dn < 0 && val) {
              double fv = sp[tt->dof].fixed_value[-dn - 1];
              b[eq - 1] += (fv * 1.) * (-val);
            }
          }
        }
      }
    }
  }
  free(Ek);
  free(kk);
  if(err) {
    for(int i = 0; i < N; i++) {
      free(rows[i].col);
      free(rows[i].val);
    }
    free(rows);
    return -1;
  }
  return rows_to_csr(rows, N, ia, ja, a, max_nnz, T);
}

long orc_assemble(const orc_mesh *m, int nspaces, const orc_space *sp,
                  int nterms, const orc_term *t, int nmax, const int *dofnum,
                  int N, int *ia, int *ja, double *a, long max_nnz, double *b)
{
  if(g_threads > 1) return assemble_threaded(m, sp, nterms, t, nmax, dofnum, N, ia, ja, a, max_nnz, b);
  row_list *rows = (row_list *)calloc((size_t)N, sizeof(row_list));
  for(int i = 0; i < N; i++) b[i] = 0.;
  (void)nspaces;
  for(int e = 0; e < m->nelem; e++) {
    for(int ti = 0; ti < nterms; ti++) {
      const orc_term *tt = &t[ti];
      if(!term_active(m, sp, tt, e)) continue;
      double Ek[MAXK][MAXK];
      int ke, kd;
      if(elem_matrix_impl(m, tt, sp[tt->equ].bf == ORC_BF_PERP_EDGE, e, Ek, &ke, &kd)) {
        for(int i = 0; i < N; i++) {
          free(rows[i].col);
          free(rows[i].val);
        }
        free(rows);
        return -1;
      }
      /* Cal_AssembleTerm_NoDt -> Dof_AssembleInMat (DofData.cpp:1464-1500) */
      for(int i = 0; i < ke; i++) {
        int sg;
        int eq_ent = space_entity(m, &sp[tt->equ], e, i, &sg);
        int eq = dofnum[(size_t)tt->equ * (nmax + 1) + eq_ent];
        if(eq <= 0) continue; /* fixed equation: nothing assembled */
        for(int j = 0; j < kd; j++) {
          double val = Ek[i][j];
          if(tt->dof < 0) { /* pseudo-DOF: fixed, value 1, time function 1 */
            if(val) b[eq - 1] += (1. * 1.) * (-val);
            continue;
          }
          int dent = space_entity(m, &sp[tt->dof], e, j, &sg);
          int dn = dofnum[(size_t)tt->dof * (nmax + 1) + dent];
          if(dn > 0) row_add(&rows[eq - 1], dn - 1, val);
          else if(dn < 0 && val) {
            double fv = sp[tt->dof].fixed_value[-dn - 1];
            b[eq - 1] += (fv * 1.) * (-val);
          }
        }
      }
    }
  }
  /* csr_format (Sparskit.cpp:2186-2208): rows by equation, columns sorted */
  long nnz = 0;
  ia[0] = 0;
  for(int i = 0; i < N; i++) {
    row_list *r = &rows[i];
    for(int p = 1; p < r->n; p++) { /* insertion sort by column */
      int c = r->col[p];
      double v = r->val[p];
      int q = p - 1;
      while(q >= 0 && r->col[q] > c) {
        r->col[q + 1] = r->col[q];
        r->val[q + 1] = r->val[q];
        q--;
      }
      r->col[q + 1] = c;
      r->val[q + 1] = v;
    }
    for(int p = 0; p < r->n; p++) {
      if(nnz < max_nnz) {
        ja[nnz] = r->col[p];
        a[nnz] = r->val[p];
      }
      nnz++;
    }
    ia[i + 1] = (int)nnz;
    free(r->col);
    free(r->val);
  }
  free(rows);
  return nnz;
}

/* ---- Post-processing ---------------------------------------------------- */

/* Pos_LocalOrIntegralQuantity (Cal_PostQuantity.cpp:29-247) for one quantity
 * term on P1 elements, with Pos_FemInterpolation's sum (Pos_FemInterpolation.cpp
 * :253-308: Val += BF * DofValue over the element's basis functions in order,
 * the DOF value being the unknown's or the constraint's, Dof_GetRealDofValue).
 * Local quantities are evaluated at each vertex (OnElementsOf); the integral
 * one sums Factor = weight * |detJ| times coef * SquNorm[{d u}] over the
 * Gauss points (Cal_AddMultValue, Cal_Value.cpp:357-437; F_SquNorm,
 * F_ExtMath.cpp:183-218). */
int orc_post(const orc_mesh *m, const orc_space *sp, const orc_post_q *q, const int *dofnum,
             const double *x, double *out, double *total)
{
  const orc_space *s = sp + q->space;
  int ns = 0;
  double tot = 0.0;
  for(int e = 0; e < m->nelem; e++) {
    if(!in_list(m->region[e], q->nregions, q->regions)) continue;
    int t = m->etype[e], tri = (t == ORC_TRI), nn = nnodes_of(t);
    double xs[4], ys[4], zs[4];
    for(int i = 0; i < nn; i++) {
      int nd = m->conn[(size_t)e * 4 + i] - 1;
      xs[i] = m->xyz[3 * nd];
      ys[i] = m->xyz[3 * nd + 1];
      zs[i] = m->xyz[3 * nd + 2];
    }
    m3 J, I;
    double d = jacobian(t, 0, xs, ys, zs, &J);
    inverse(!tri, d, &J, &I);
    int edge = s->bf == ORC_BF_EDGE, perp = s->bf == ORC_BF_PERP_EDGE;
    int nb = in_list(m->region[e], s->nsupport, s->support) ? (edge ? 6 : nn) : 0;
    double dv[6];
    int has[6], neg[6];
    for(int j = 0; j < nb; j++) {
      int sign;
      int ent = space_entity(m, s, e, j, &sign);
      int c = dofnum[ent];
      neg[j] = sign < 0;
      has[j] = c != 0;
      dv[j] = c > 0 ? x[c - 1] : (c < 0 ? s->fixed_value[-c - 1] : 0.0);
    }
    double D[3] = {0.0, 0.0, 0.0};
    if(q->kind != ORC_PQ_VALUE) {
      for(int j = 0; j < nb; j++) {
        if(!has[j]) continue;
        double r[3], b[3];
        if(edge) {
          bf_curledge_tet(j, r);
          if(neg[j]) { r[0] = -r[0]; r[1] = -r[1]; r[2] = -r[2]; }
          form2(&J, d, r, b);
        }
        else if(perp) {
          bf_curlperp(t, j, r);
          form2p(&J, d, r, b);
        }
        else {
          bf_gradnode(t, j, r);
          form1(&I, r, b);
        }
        D[0] += b[0] * dv[j];
        D[1] += b[1] * dv[j];
        D[2] += b[2] * dv[j];
      }
    }
    const double *cf = q->coef ? q->coef + (size_t)m->region[e] * q->ncoef : NULL;
    if(q->kind == ORC_PQ_INT_COEF_D2) {
      double tmp = cf[0] * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
      double T = 0.0;
      for(int g = 0; g < q->npoints; g++) {
        double u, v, w, wt;
        gauss_point(t, q->npoints, g, &u, &v, &w, &wt);
        T = T + tmp * (wt * fabs(d));
      }
      if(out) out[ns] = T;
      tot = tot + T;
      ns++;
      continue;
    }
    double *o = out ? out + (size_t)ns * 12 : NULL;
    ns++;
    if(!o) continue;
    for(int i = 0; i < 12; i++) o[i] = 0.0;
    if(q->kind != ORC_PQ_VALUE) {
      double R[3] = {D[0], D[1], D[2]};
      if(q->kind == ORC_PQ_COEF_D) coef_product(q->ncoef, cf, D, R);
      for(int i = 0; i < nn; i++)
        for(int c = 0; c < 3; c++) o[i * 3 + c] = R[c];
      continue;
    }
    for(int i = 0; i < nn; i++) {
      double u = i == 1 ? 1.0 : 0.0, v = i == 2 ? 1.0 : 0.0, w = i == 3 ? 1.0 : 0.0;
      double V[3] = {0.0, 0.0, 0.0};
      for(int j = 0; j < nb; j++) {
        if(!has[j]) continue;
        if(edge) {
          double r[3], b[3];
          bf_edge_tet(j, u, v, w, r);
          if(neg[j]) { r[0] = -r[0]; r[1] = -r[1]; r[2] = -r[2]; }
          form1(&I, r, b);
          V[0] += b[0] * dv[j];
          V[1] += b[1] * dv[j];
          V[2] += b[2] * dv[j];
        }
        else if(perp) {
          double r[3] = {0.0, 0.0, bf_node(t, j, u, v, w)}, b[3];
          form1p(&J, r, b);
          V[0] += b[0] * dv[j];
          V[1] += b[1] * dv[j];
          V[2] += b[2] * dv[j];
        }
        else
          V[0] += bf_node(t, j, u, v, w) * dv[j];
      }
      o[i * 3 + 0] = V[0];
      o[i * 3 + 1] = V[1];
      o[i * 3 + 2] = V[2];
    }
  }
  if(total) *total = tot;
  return ns;
}
