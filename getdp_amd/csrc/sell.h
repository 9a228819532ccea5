// Sliced-ELLPACK storage (slices of 64 rows = one wavefront) for SpMV.
//
// Row r lives in slice s = r / 64 at lane l = r % 64.  Its entries, in column
// order, are taken in pairs: entries 2q and 2q+1 are stored side by side at
// sp[s] + 128 q + 2 l (+1), so that one 16-byte load per lane fetches a pair
// and one load instruction of a wave reads 64 consecutive pairs (1 KiB of
// FP64 values).  A slice is as wide as its longest row, rounded up to even;
// shorter rows are padded with zero values whose column is the row itself.
// One lane computes one row and sums it sequentially in column order, the
// order of the reference's amux (matvec.f:35-76).
#pragma once

#include "gdp_device.h"

namespace gdp {

constexpr int kSlice = 64;

typedef double sell_d2 __attribute__((ext_vector_type(2)));
typedef float sell_f2 __attribute__((ext_vector_type(2)));
typedef int sell_i2 __attribute__((ext_vector_type(2)));
template <typename T> struct SellPair;
template <> struct SellPair<double> { typedef sell_d2 type; };
template <> struct SellPair<float> { typedef sell_f2 type; };

// Pattern from a host CSR; values copied from a device CSR with the same
// pattern (also usable alone to refresh values).
int sell_build(const std::vector<int> &h_ia, const DevCSR &A, DevSell &S);
int sell_copy_values(const DevCSR &A, DevSell &S, bool also_f32);
int sell_spmv(const DevSell &S, const double *x, double *y);

// Compressed columns.  When every pair group (the 64 lanes x 2 entries of one
// slice at one pair position q) spans fewer than 65536 columns, the group
// keeps one 32-bit base, cbase[sp[s] / 128 + q], and each lane one 32-bit word
// holding the two 16-bit offsets from it, c16[sp[s] / 2 + 64 q + l]: an FP64
// entry then costs 10 bytes of index + value instead of 12.  Bandwidth-limited
// orderings (RCMK) fit; when one group does not, the 32-bit columns stay.
template <typename TA> struct SellOp {
  int n = 0;
  const long *sp = nullptr;
  const int *w = nullptr;
  const int *col = nullptr;         // 32-bit columns, or null when compressed
  const unsigned *c16 = nullptr;    // compressed columns (null: 32-bit)
  const int *cbase = nullptr;
  const TA *val = nullptr;
};

template <typename TA> SellOp<TA> sell_op(const DevSell &S);
template <> inline SellOp<double> sell_op<double>(const DevSell &S)
{
  return SellOp<double>{S.n, S.sp, S.w, S.col, S.c16, S.cbase, S.val};
}
template <> inline SellOp<float> sell_op<float>(const DevSell &S)
{
  return SellOp<float>{S.n, S.sp, S.w, S.col, S.c16, S.cbase, S.val32};
}

// y = A x, one row per lane; U pairs of entries in flight per lane.  The
// slice index is wave-uniform (a wave is one slice), so the slice offsets and
// the group bases are scalar loads.
template <typename TA, typename TV, int U = 2>
__device__ inline void sell_row(int row, const SellOp<TA> &A, const TV *__restrict__ x, double &sum)
{
#pragma clang fp contract(off)
  typedef typename SellPair<TA>::type VP;
  const int s = __builtin_amdgcn_readfirstlane(row >> 6), lane = row & 63;
  const long sps = A.sp[s];
  const int width = A.w[s];
  const VP *vp = (const VP *)(A.val + sps + 2 * lane);
  sum = 0.;
  const int nq = width >> 1;
  int q = 0;
  if(A.c16) {
    const unsigned *cw = A.c16 + (sps >> 1) + lane;
    const int *cb = A.cbase + (sps >> 7);
    for(; q + U <= nq; q += U) {
      VP v[U];
      unsigned c[U];
      double xa[U], xb[U];
#pragma unroll
      for(int u = 0; u < U; u++) {
        v[u] = __builtin_nontemporal_load(vp + (long)(q + u) * kSlice);
        c[u] = __builtin_nontemporal_load(cw + (long)(q + u) * kSlice);
      }
#pragma unroll
      for(int u = 0; u < U; u++) {
        const int b = cb[q + u];
        xa[u] = (double)x[b + (int)(c[u] & 0xffffu)];
        xb[u] = (double)x[b + (int)(c[u] >> 16)];
      }
#pragma unroll
      for(int u = 0; u < U; u++) {
        sum = sum + (double)v[u].x * xa[u];
        sum = sum + (double)v[u].y * xb[u];
      }
    }
    for(; q < nq; q++) {
      const VP v = __builtin_nontemporal_load(vp + (long)q * kSlice);
      const unsigned c = __builtin_nontemporal_load(cw + (long)q * kSlice);
      const int b = cb[q];
      sum = sum + (double)v.x * (double)x[b + (int)(c & 0xffffu)];
      sum = sum + (double)v.y * (double)x[b + (int)(c >> 16)];
    }
    return;
  }
  const sell_i2 *cp = (const sell_i2 *)(A.col + sps + 2 * lane);
  for(; q + U <= nq; q += U) {
    VP v[U];
    sell_i2 c[U];
    double xa[U], xb[U];
#pragma unroll
    for(int u = 0; u < U; u++) {
      v[u] = __builtin_nontemporal_load(vp + (long)(q + u) * kSlice);
      c[u] = __builtin_nontemporal_load(cp + (long)(q + u) * kSlice);
    }
#pragma unroll
    for(int u = 0; u < U; u++) {
      xa[u] = (double)x[c[u].x];
      xb[u] = (double)x[c[u].y];
    }
#pragma unroll
    for(int u = 0; u < U; u++) {
      sum = sum + (double)v[u].x * xa[u];
      sum = sum + (double)v[u].y * xb[u];
    }
  }
  for(; q < nq; q++) {
    const VP v = __builtin_nontemporal_load(vp + (long)q * kSlice);
    const sell_i2 c = __builtin_nontemporal_load(cp + (long)q * kSlice);
    sum = sum + (double)v.x * (double)x[c.x];
    sum = sum + (double)v.y * (double)x[c.y];
  }
}

}  // namespace gdp
