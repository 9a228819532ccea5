// ILUT and ILUTP factorisation on the host (Sparskit ilut.F: ilut :52-396,
// ilutp :397-806, as Sparskit.cpp:375-390 calls them with mbloc = n).
//
// Row i of the factor is formed from row i of A by eliminating its strictly
// lower entries in ascending column order with the rows of U already built,
// keeping a multiplier only when |multiplier| > droptol; then the lfil
// largest multipliers form row i of L, the upper entries with
// |w| > droptol * (mean |a_ij| of the row) are cut to the lfil - 1 largest
// (partial quick-split, whose output order is kept: it is the order in which
// the triangular sweeps accumulate), and ILUTP may swap the diagonal column
// with the largest upper entry when that entry exceeds the diagonal by the
// factor 1/permtol.  A zero pivot is replaced by (1e-4 + droptol) * mean |a_ij|
// (ilut writes the 1e-4 as a single-precision literal).  Fill is unbounded
// here: the reference grows its work arrays by half and factors again when
// they overflow, which gives the same factors.
//
// The factorisation is sequential by nature (which entries a row keeps depends
// on every earlier row's kept entries and on the values); it runs once per
// matrix, the sweeps that apply the factors every iteration run on the device.
#include <cmath>
#include <utility>

#include "gdp_internal.h"

namespace gdp {

namespace {

// qsplit (ilut.F:2403-2460) on v[0..m), idx[0..m): reorders so that the ncut
// entries of largest magnitude come first.
void qsplit(double *v, int *idx, int m, int ncut)
{
  int first = 0, last = m - 1;
  const int cut = ncut - 1;
  if(cut < first || cut > last) return;
  for(;;) {
    int mid = first;
    const double key = std::fabs(v[mid]);
    for(int j = first + 1; j <= last; j++) {
      if(std::fabs(v[j]) > key) {
        ++mid;
        std::swap(v[mid], v[j]);
        std::swap(idx[mid], idx[j]);
      }
    }
    std::swap(v[mid], v[first]);
    std::swap(idx[mid], idx[first]);
    if(mid == cut) return;
    if(mid > cut) last = mid - 1;
    else first = mid + 1;
  }
}

}  // namespace

template <typename T>
int ilut_factor(int n, const int *ia, const int *ja, const double *a, int lfil, double droptol, double permtol,
                bool pivot, IlutRows<T> &f)
{
  if(lfil < 0) return set_error(GDP_E_ARG, "Illegal value of nb_fill in ILU");
  f.ptr.assign(1, 0);
  f.ptr.reserve((size_t)n + 1);
  f.dpos.assign(n, 0);
  f.col.clear();
  f.val.clear();
  f.col.reserve((size_t)n * (2 * (size_t)lfil + 1));
  f.val.reserve(f.col.capacity());
  // oldcol[p]: input column now at position p; newpos[c]: position of input column c
  std::vector<int> oldcol(n), newpos(n);
  for(int c = 0; c < n; c++) oldcol[c] = newpos[c] = c;
  bool ident = true;  // no column swapped yet: positions are the input columns
  // working row dense by column position (w, present where at[c] = 1); the
  // lower columns pending elimination are bits of lbits, the upper columns
  // are listed in order of appearance in ju (the diagonal first), and the
  // kept multipliers go to (wl, jl) in elimination order.
  //
  // ilut.F picks the next row to eliminate by scanning the pending lower
  // columns for the smallest, O(lenl^2) per row.  Here they are visited in
  // ascending order of their bits: a row jrow's upper entries sit at
  // positions above jrow (ILUTP's later swaps exchange positions at or above
  // the row being formed), so every fill-in lands ahead of the scan and the
  // columns are eliminated in the order the reference eliminates them, each
  // entry receiving the same updates in the same order.  The rows of A hold
  // distinct columns (an assembled CSR pattern).
  std::vector<double> w(n), wl(n), wu(n + 1);
  std::vector<int> jl(n), ju(n + 1);
  std::vector<unsigned char> at(n, 0);
  std::vector<unsigned long long> lbits((size_t)n / 64 + 1, 0ull);
  for(int i = 0; i < n; i++) {
    const int j1 = ia[i], j2 = ia[i + 1];
    double tnorm = 0.0;
    for(int k = j1; k < j2; k++) tnorm = tnorm + std::fabs(a[k]);
    if(tnorm == 0.0) return set_error(GDP_E_ZERO_ROW, "Zero row encountered in ILU (row %d)", i + 1);
    tnorm = tnorm / (double)(j2 - j1);
    int lenu = 1, lo = i;
    ju[0] = i;
    w[i] = 0.0;
    at[i] = 1;
    for(int k = j1; k < j2; k++) {
      const int c = newpos[ja[k]];
      w[c] = a[k];
      at[c] = 1;
      if(c < i) {
        lbits[(size_t)c >> 6] |= 1ull << (c & 63);
        lo = std::min(lo, c);
      }
      else if(c > i)
        ju[lenu++] = c;
    }
    // elimination, smallest remaining lower column first
    int kept = 0;
    const size_t wend = (size_t)i >> 6;
    for(size_t wd = (size_t)lo >> 6; wd <= wend;) {
      const unsigned long long bits = lbits[wd];
      if(!bits) {
        wd++;
        continue;
      }
      const int jrow = (int)(wd << 6) + __builtin_ctzll(bits);
      lbits[wd] = bits & (bits - 1);
      at[jrow] = 0;
      const double fact = w[jrow] * (double)f.val[f.dpos[jrow]];
      if(std::fabs(fact) <= droptol) continue;
      const int k1 = f.dpos[jrow] + 1, k2 = f.ptr[jrow + 1];
      const int *fc = f.col.data();
      const T *fv = f.val.data();
      for(int kk = k1; kk < k2; kk++) {
        const double sv = fact * (double)fv[kk];
        const int c = ident ? fc[kk] : newpos[fc[kk]];
        if(at[c])
          w[c] = w[c] - sv;
        else {
          w[c] = -sv;
          at[c] = 1;
          if(c < i) lbits[(size_t)c >> 6] |= 1ull << (c & 63);
          else ju[lenu++] = c;
        }
      }
      wl[kept] = fact;
      jl[kept] = jrow;
      kept++;
    }
    for(int k = 0; k < lenu; k++) {
      wu[k] = w[ju[k]];
      at[ju[k]] = 0;
    }
    // row i of L: the lfil largest multipliers
    int len = std::min(kept, lfil);
    qsplit(wl.data(), jl.data(), kept, len);
    for(int k = 0; k < len; k++) {
      f.col.push_back(oldcol[jl[k]]);
      f.val.push_back((T)wl[k]);
    }
    f.dpos[i] = (int)f.col.size();
    f.col.push_back(0);  // pivot slot, filled below
    f.val.push_back((T)0);
    // row i of U: drop, keep the lfil - 1 largest off-diagonal entries
    len = 0;
    for(int k = 1; k < lenu; k++)
      if(std::fabs(wu[k]) > droptol * tnorm) {
        len++;
        wu[len] = wu[k];
        ju[len] = ju[k];
      }
    lenu = len + 1;
    len = std::min(lenu, lfil);
    qsplit(wu.data() + 1, ju.data() + 1, lenu - 1, len);
    if(pivot) {
      int imax = 0;
      double xmax = std::fabs(wu[0]);
      const double xmax0 = xmax;
      for(int k = 1; k < len; k++) {
        const double t = std::fabs(wu[k]);
        if(t > xmax && t * permtol > xmax0) {
          imax = k;
          xmax = t;
        }
      }
      std::swap(wu[0], wu[imax]);
      const int j = ju[imax];
      if(j != i) ident = false;
      std::swap(oldcol[i], oldcol[j]);
      newpos[oldcol[i]] = i;
      newpos[oldcol[j]] = j;
    }
    for(int k = 1; k < len; k++) {
      f.col.push_back(oldcol[ju[k]]);
      f.val.push_back((T)wu[k]);
    }
    if(wu[0] == 0.0) wu[0] = (pivot ? (1.0e-4 + droptol) : ((double)0.0001f + droptol)) * tnorm;
    f.col[f.dpos[i]] = oldcol[i];
    f.val[f.dpos[i]] = (T)(1.0 / wu[0]);
    f.ptr.push_back((int)f.col.size());
    if(f.col.size() > (size_t)0x7fffffff) return set_error(GDP_E_NOMEM, "ILU factors exceed 2^31 entries");
  }
  // columns in the final numbering
  if(pivot) {
    for(int &c : f.col) c = newpos[c];
    f.newpos.swap(newpos);
  }
  else
    f.newpos.clear();
  return GDP_OK;
}

template int ilut_factor<float>(int, const int *, const int *, const double *, int, double, double, bool,
                                IlutRows<float> &);
template int ilut_factor<double>(int, const int *, const int *, const double *, int, double, double, bool,
                                 IlutRows<double> &);

}  // namespace gdp

namespace {
template <typename T>
int export_rows(int n, const gdp::IlutRows<T> &f, long capacity, int *ptr, int *dpos, int *col, double *val,
                int *newpos, long *nnz_out)
{
  const long nz = (long)f.col.size();
  *nnz_out = nz;
  for(int i = 0; i <= n; i++) ptr[i] = f.ptr[i];
  if(nz > capacity) return gdp::set_error(GDP_E_NOMEM, "gdp_ilut: %ld entries exceed capacity %ld", nz, capacity);
  for(int i = 0; i < n; i++) {
    dpos[i] = f.dpos[i];
    newpos[i] = f.newpos.empty() ? i : f.newpos[i];
  }
  for(long k = 0; k < nz; k++) {
    col[k] = f.col[k];
    val[k] = (double)f.val[k];
  }
  return GDP_OK;
}
}  // namespace

extern "C" int gdp_ilut(int n, const int *ia, const int *ja, const double *a, int lfil, double droptol,
                        double permtol, int pivot, int f64, long capacity, int *ptr, int *dpos, int *col,
                        double *val, int *newpos, long *nnz_out)
{
  if(n < 0 || !ia || !ja || !a || !ptr || !dpos || !newpos || !nnz_out || capacity < 0 ||
     (capacity > 0 && (!col || !val)))
    return gdp::set_error(GDP_E_ARG, "gdp_ilut: bad arguments");
  if(f64) {
    gdp::IlutRows<double> f;
    int r = gdp::ilut_factor(n, ia, ja, a, lfil, droptol, permtol, pivot != 0, f);
    return r ? r : export_rows(n, f, capacity, ptr, dpos, col, val, newpos, nnz_out);
  }
  gdp::IlutRows<float> f;
  int r = gdp::ilut_factor(n, ia, ja, a, lfil, droptol, permtol, pivot != 0, f);
  return r ? r : export_rows(n, f, capacity, ptr, dpos, col, val, newpos, nnz_out);
}
