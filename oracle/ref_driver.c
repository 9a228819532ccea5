/*
 * ref_driver.c -- test infrastructure only (never shipped, never measured as
 * the product).  Drives the reference's own Fortran Sparskit kernels,
 * compiled from contrib/Sparskit/{iters.f,ilut.F,matvec.f,blas1.f,blassm.f,
 * unary.f,cmkreord.f,reordering.f} into oracle/_ref/, through the same
 * sequence of calls that the reference's C++ glue performs in
 * contrib/Sparskit/Sparskit.cpp:108-634 (solve_matrix).  That glue cannot be
 * compiled on its own (it needs GetDP's generated config header and its
 * List/Message utilities), so this file restates its control flow; every
 * arithmetic kernel it calls is the reference's.
 *
 * All matrices are 1-based CSR (Sparskit's convention).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef REF_ILU_DOUBLE
typedef double sscalar;
#else
typedef float sscalar; /* CMakeLists.txt:552 forces -DHAVE_ILU_FLOAT */
#endif

extern void amux_(int *n, double *x, double *y, double *a, int *ja, int *ia);
extern void atmux_(int *n, double *x, double *y, double *a, int *ja, int *ia);
extern void ilu0_(int *n, double *a, int *ja, int *ia, sscalar *alu, int *jlu,
                  int *ju, int *iw, int *ierr);
extern void lusol_(int *n, double *y, double *x, sscalar *alu, int *jlu, int *ju);
extern void ilut_(int *n, double *a, int *ja, int *ia, int *lfil, double *droptol,
                  sscalar *alu, int *jlu, int *ju, int *iwk, double *w, int *jw, int *ierr);
extern void ilutp_(int *n, double *a, int *ja, int *ia, int *lfil, double *droptol,
                   double *permtol, int *mbloc, sscalar *alu, int *jlu, int *ju, int *iwk,
                   double *w, int *jw, int *iperm, int *ierr);
extern void lutsol_(int *n, double *y, double *x, sscalar *alu, int *jlu,
                    int *ju);
extern void cg_(int *n, double *rhs, double *sol, int *ipar, double *fpar,
                double *w);
extern void bcgstab_(int *n, double *rhs, double *sol, int *ipar, double *fpar,
                     double *w);
extern void tfqmr_(int *n, double *rhs, double *sol, int *ipar, double *fpar,
                   double *w);
extern void gmres_(int *n, double *rhs, double *sol, int *ipar, double *fpar,
                   double *w);
extern void fgmres_(int *n, double *rhs, double *sol, int *ipar, double *fpar,
                    double *w);
extern void cmkreord_(int *n, double *a, int *ja, int *ia, double *a0, int *ja0,
                      int *ia0, int *init, int *iperm, int *mask, int *maskval,
                      int *nlev, int *riord, int *levels);
extern void sortcol_(int *n, double *a, int *ja, int *ia, int *iw, double *rw);
extern void getdia_(int *nrow, int *ncol, int *job, double *a, int *ja, int *ia,
                    int *len, double *diag, int *idiag, int *ioff);
extern void diamua_(int *nrow, int *job, double *a, int *ja, int *ia,
                    double *diag, double *b, int *jb, int *ib);
extern void amudia_(int *nrow, int *job, double *a, int *ja, int *ia,
                    double *diag, double *b, int *jb, int *ib);
extern double dnrm2_(int *n, double *x, int *incx);
extern double ddot_(int *n, double *x, int *incx, double *y, int *incy);

int ref_sscalar_size(void) { return (int)sizeof(sscalar); }

/* Sparskit.h algorithm / preconditioner codes (contrib/Sparskit/Sparskit.h). */
enum { A_CG = 1, A_BCGSTAB = 5, A_TFQMR = 6, A_GMRES = 8, A_FGMRES = 9 };
enum { P_NONE = 0, P_ILUT = 1, P_ILUTP = 2, P_ILU0 = 6, P_DIAGONAL = 8 };

typedef struct {
  int algorithm, preconditioner, precond_position, krylov, maxits;
  int renumbering, scaling;
  double rtol;
  int nb_fill;
  double droptol, permtol;
} ref_params;

typedef struct {
  int its;          /* ipar(7) at exit: matrix-vector products performed */
  int status;       /* ipar(1) at exit */
  int ilu_ierr;     /* ilu0 error code */
  double res_est;   /* fpar(6) */
  double res1;      /* fpar(5) captured at its==1 (log reference) */
  double true_res;  /* dnrm2(A*sol - rhs) in the permuted/scaled system */
} ref_info;

/* The permutation produced by cmkreord+sortcol, exposed for RCMK parity. */
int ref_rcmk(int n, double *a, int *ja, int *ia, double *a_out, int *ja_out,
             int *ia_out, int *permr, int *rpermr)
{
  int nnz = ia[n] - 1;
  int *mask = (int *)malloc(sizeof(int) * (size_t)(nnz > n ? nnz : n));
  int *levels = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)nnz);
  int init = 1, maskval = 1, nlev = 1;
  for(int i = 0; i < n; i++) permr[i] = rpermr[i] = i + 1;
  cmkreord_(&n, a, ja, ia, a_out, ja_out, ia_out, &init, permr, mask, &maskval,
            &nlev, rpermr, levels);
  sortcol_(&n, a_out, ja_out, ia_out, mask, w);
  free(w);
  free(mask);
  free(levels);
  return nlev;
}

int ref_ilu0(int n, double *a, int *ja, int *ia, sscalar *alu, int *jlu, int *ju)
{
  int ierr = 0;
  int *iw = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  ilu0_(&n, a, ja, ia, alu, jlu, ju, iw, &ierr);
  free(iw);
  return ierr;
}

/* ilut / ilutp on a 1-based CSR matrix (ja permuted in place by ilutp);
 * iwk entries of room; iperm (2n) for ilutp. */
int ref_ilut(int n, double *a, int *ja, int *ia, int lfil, double droptol, double permtol,
             int pivot, int iwk, sscalar *alu, int *jlu, int *ju, int *iperm)
{
  int ierr = 0;
  double *w = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  int *jw = (int *)malloc(sizeof(int) * 2 * (size_t)(n + 1));
  if(pivot) {
    int mbloc = n;
    ilutp_(&n, a, ja, ia, &lfil, &droptol, &permtol, &mbloc, alu, jlu, ju, &iwk, w, jw, iperm,
           &ierr);
  }
  else
    ilut_(&n, a, ja, ia, &lfil, &droptol, alu, jlu, ju, &iwk, w, jw, &ierr);
  free(w);
  free(jw);
  return ierr;
}

void ref_lusol(int n, double *y, double *x, sscalar *alu, int *jlu, int *ju)
{
  lusol_(&n, y, x, alu, jlu, ju);
}

void ref_amux(int n, double *x, double *y, double *a, int *ja, int *ia)
{
  amux_(&n, x, y, a, ja, ia);
}

/* solve_matrix (Sparskit.cpp:108-634) for the SPARSE format, the
 * preconditioners NONE / ILU0 / DIAGONAL, the algorithms CG / BCGSTAB / TFQMR /
 * GMRES / FGMRES,
 * renumbering NONE / RCMK and scaling NONE / DIAG.  a, ja, ia are modified in
 * place by scaling, exactly as the reference modifies its matrix. */
int ref_solve(int n, double *a, int *ja, int *ia, double *b, double *x,
              const ref_params *p, ref_info *info)
{
  int nnz = ia[n] - 1;
  double *rowscal = NULL;
  memset(info, 0, sizeof(*info));

  if(p->scaling == 1) { /* Sparskit.cpp:975-995 */
    int job0 = 0, job1 = 1, ioff = 0, len;
    int *idiag = (int *)malloc(sizeof(int) * (size_t)n);
    rowscal = (double *)malloc(sizeof(double) * (size_t)n);
    getdia_(&n, &n, &job0, a, ja, ia, &len, rowscal, idiag, &ioff);
    free(idiag);
    for(int i = 0; i < n; i++) {
      if(rowscal[i]) rowscal[i] = 1. / sqrt(fabs(rowscal[i]));
      else { free(rowscal); rowscal = NULL; break; }
    }
    if(rowscal) {
      diamua_(&n, &job1, a, ja, ia, rowscal, a, ja, ia);
      amudia_(&n, &job1, a, ja, ia, rowscal, a, ja, ia);
      for(int i = 0; i < n; i++) b[i] *= rowscal[i];
    }
  }

  for(int i = 1; i < n; i++)
    if(ia[i] - ia[i - 1] <= 0) return -100; /* "Zero row in matrix" */

  double *rhs = (double *)malloc(sizeof(double) * (size_t)n);
  double *sol = (double *)calloc((size_t)n, sizeof(double));
  int *permr = (int *)malloc(sizeof(int) * (size_t)n);
  int *rpermr = (int *)malloc(sizeof(int) * (size_t)n);
  double *A = a;
  int *JA = ja, *IA = ia;
  double *a_r = NULL;
  int *ja_r = NULL, *ia_r = NULL;
  for(int i = 0; i < n; i++) permr[i] = rpermr[i] = i + 1;
  if(p->renumbering == 1) {
    a_r = (double *)malloc(sizeof(double) * (size_t)nnz);
    ja_r = (int *)malloc(sizeof(int) * (size_t)nnz);
    ia_r = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    ref_rcmk(n, a, ja, ia, a_r, ja_r, ia_r, permr, rpermr);
    A = a_r;
    JA = ja_r;
    IA = ia_r;
  }

  sscalar *alu = NULL;
  int *jlu = NULL, *ju = NULL, *permp = NULL;
  if(p->preconditioner == P_ILUT || p->preconditioner == P_ILUTP) {
    /* Sparskit.cpp:312-317, 375-390, 435-446: first guess, grow by half on
     * -2/-3 and factor again */
    int nnz_ilu = 2 * (n + 1) * (p->nb_fill + 1);
    permp = (int *)malloc(sizeof(int) * 2 * (size_t)n);
    int *ja_keep = (int *)malloc(sizeof(int) * (size_t)nnz);
    memcpy(ja_keep, JA, sizeof(int) * (size_t)nnz);
    for(;;) {
      alu = (sscalar *)malloc(sizeof(sscalar) * (size_t)nnz_ilu);
      jlu = (int *)malloc(sizeof(int) * (size_t)nnz_ilu);
      ju = (int *)malloc(sizeof(int) * (size_t)(n + 1));
      memcpy(JA, ja_keep, sizeof(int) * (size_t)nnz);
      int ierr = ref_ilut(n, A, JA, IA, p->nb_fill, p->droptol, p->permtol,
                          p->preconditioner == P_ILUTP, nnz_ilu, alu, jlu, ju, permp);
      if(ierr == -2 || ierr == -3) {
        nnz_ilu += nnz_ilu / 2;
        free(alu);
        free(jlu);
        free(ju);
        continue;
      }
      info->ilu_ierr = ierr;
      break;
    }
    free(ja_keep);
    if(p->preconditioner == P_ILUT)
      for(int i = 0; i < n; i++) permp[n + i] = i + 1;
    if(info->ilu_ierr) goto done;
  }
  else if(p->preconditioner == P_ILU0) {
    int nnz_ilu = nnz + n + 1;
    alu = (sscalar *)malloc(sizeof(sscalar) * (size_t)nnz_ilu);
    jlu = (int *)malloc(sizeof(int) * (size_t)nnz_ilu);
    ju = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    info->ilu_ierr = ref_ilu0(n, A, JA, IA, alu, jlu, ju);
    if(info->ilu_ierr) goto done;
  }
  else if(p->preconditioner == P_DIAGONAL) { /* Sparskit.cpp:338-354 */
    alu = (sscalar *)malloc(sizeof(sscalar) * (size_t)(n + 1));
    jlu = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    ju = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    for(int i = 0; i < n; i++) {
      alu[i] = 1.0;
      jlu[i] = n + 2;
      ju[i] = n + 2;
    }
    alu[n] = 0.0;
    jlu[n] = n + 2;
    ju[n] = 0;
  }

  for(int i = 0; i < n; i++) rhs[i] = b[rpermr[i] - 1];

  {
    int ipar[17];
    double fpar[17];
    memset(ipar, 0, sizeof(ipar));
    memset(fpar, 0, sizeof(fpar));
    ipar[1] = 0;
    ipar[2] = (p->preconditioner == P_NONE) ? 0 : p->precond_position;
    ipar[3] = 1;
    ipar[5] = p->krylov;
    ipar[6] = p->maxits;
    fpar[1] = p->rtol;
    fpar[2] = 0.0;
    fpar[11] = 0.0;
    int m = ipar[5] <= 1 ? 15 : ipar[5];
    switch(p->algorithm) {
    case A_CG: ipar[4] = 5 * n; break;
    case A_BCGSTAB: ipar[4] = 8 * n; break;
    case A_TFQMR: ipar[4] = 11 * n; break;
    case A_GMRES: ipar[4] = (n + 3) * (m + 2) + (m + 1) * m / 2; break;
    case A_FGMRES:
      ipar[4] = 2 * n * (m + 1) + (m + 1) * m / 2 + 3 * m + 2;
      break;
    default: info->status = -200; goto done;
    }
    double *w = (double *)malloc(sizeof(double) * (size_t)ipar[4]);
    int its = 0, end = 0;
    double res1 = 1.;
    for(;;) {
      switch(p->algorithm) {
      case A_CG: cg_(&n, rhs, sol, &ipar[1], &fpar[1], w); break;
      case A_BCGSTAB: bcgstab_(&n, rhs, sol, &ipar[1], &fpar[1], w); break;
      case A_TFQMR: tfqmr_(&n, rhs, sol, &ipar[1], &fpar[1], w); break;
      case A_GMRES: gmres_(&n, rhs, sol, &ipar[1], &fpar[1], w); break;
      case A_FGMRES: fgmres_(&n, rhs, sol, &ipar[1], &fpar[1], w); break;
      }
      if(ipar[7] != its) its = ipar[7];
      if(its == 1) res1 = fpar[5];
      switch(ipar[1]) {
      case 1: amux_(&n, &w[ipar[8] - 1], &w[ipar[9] - 1], A, JA, IA); break;
      case 2: atmux_(&n, &w[ipar[8] - 1], &w[ipar[9] - 1], A, JA, IA); break;
      case 3:
      case 5: lusol_(&n, &w[ipar[8] - 1], &w[ipar[9] - 1], alu, jlu, ju); break;
      case 4:
      case 6: lutsol_(&n, &w[ipar[8] - 1], &w[ipar[9] - 1], alu, jlu, ju); break;
      default: end = 1; break;
      }
      if(end) break;
    }
    info->its = ipar[7];
    info->status = ipar[1];
    info->res_est = fpar[6];
    info->res1 = res1;
    double *w2 = (double *)malloc(sizeof(double) * (size_t)n);
    amux_(&n, sol, w2, A, JA, IA);
    for(int i = 0; i < n; i++) w2[i] -= rhs[i];
    int one = 1;
    info->true_res = dnrm2_(&n, w2, &one);
    free(w2);
    free(w);
  }
  for(int i = 0; i < n; i++) /* Sparskit.cpp:597-602 */
    x[i] = permp ? sol[permp[n + permr[i] - 1] - 1] : sol[permr[i] - 1];
  if(rowscal)
    for(int i = 0; i < n; i++) x[i] *= rowscal[i];

done:
  free(rhs);
  free(sol);
  free(permr);
  free(rpermr);
  free(a_r);
  free(ja_r);
  free(ia_r);
  free(alu);
  free(jlu);
  free(ju);
  free(permp);
  free(rowscal);
  return info->ilu_ierr ? info->ilu_ierr : 0;
}
