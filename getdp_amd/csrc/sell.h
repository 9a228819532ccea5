// Sliced-ELLPACK storage (slices of 64 rows = one wavefront) for SpMV.
//
// Row r lives in slice s = r / 64 at lane l = r % 64; its k-th entry (in
// column order) is stored at sp[s] + k*64 + l.  A slice is as wide as its
// longest row; shorter rows are padded with zero values whose column is the
// row itself.  One lane computes one row: every load instruction of a wave
// reads 64 consecutive values (512 bytes) and 64 consecutive column indices,
// and each row is summed sequentially in column order, the order of the
// reference's amux (matvec.f:35-76).
#pragma once

#include "gdp_device.h"

namespace gdp {

constexpr int kSlice = 64;

// Pattern from a host CSR; values copied from a device CSR with the same
// pattern (also usable alone to refresh values).
int sell_build(const std::vector<int> &h_ia, const DevCSR &A, DevSell &S);
int sell_copy_values(const DevCSR &A, DevSell &S, bool also_f32);
int sell_spmv(const DevSell &S, const double *x, double *y);

// y = A x, one row per lane; U independent entries in flight per lane.
template <typename TA, typename TV, int U = 4>
__device__ inline void sell_row(int row, const long *__restrict__ sp, const int *__restrict__ w,
                                const int *__restrict__ col, const TA *__restrict__ val,
                                const TV *__restrict__ x, double &sum)
{
#pragma clang fp contract(off)
  const int s = row >> 6, lane = row & 63;
  const long base = sp[s] + lane;
  const int width = w[s];
  sum = 0.;
  int k = 0;
  for(; k + U <= width; k += U) {
    TA v[U];
    int c[U];
    double xv[U];
#pragma unroll
    for(int u = 0; u < U; u++) {
      v[u] = __builtin_nontemporal_load(val + base + (long)(k + u) * kSlice);
      c[u] = __builtin_nontemporal_load(col + base + (long)(k + u) * kSlice);
    }
#pragma unroll
    for(int u = 0; u < U; u++) xv[u] = (double)x[c[u]];
#pragma unroll
    for(int u = 0; u < U; u++) sum = sum + (double)v[u] * xv[u];
  }
  for(; k < width; k++) {
    TA v = __builtin_nontemporal_load(val + base + (long)k * kSlice);
    int c = __builtin_nontemporal_load(col + base + (long)k * kSlice);
    sum = sum + (double)v * (double)x[c];
  }
}

}  // namespace gdp
