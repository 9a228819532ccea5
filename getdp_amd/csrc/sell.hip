// SELL-64 construction and the plain FP64 SpMV entry (layout in sell.h).
#include <cstdlib>

#include "sell.h"

namespace gdp {

namespace {

// thread per row: copy the row's CSR entries into its slice column, pad
template <bool F32>
__global__ void __launch_bounds__(256) k_sell_fill(int n, const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const double *__restrict__ a,
                                                   const long *__restrict__ sp,
                                                   const int *__restrict__ w, int *__restrict__ col,
                                                   double *__restrict__ val, float *__restrict__ val32,
                                                   bool fill_cols)
{
  int r = blockIdx.x * 256 + threadIdx.x;
  int nrows_pad = ((n + kSlice - 1) / kSlice) * kSlice;
  if(r >= nrows_pad) return;
  const int s = r >> 6, lane = r & 63;
  const long base = sp[s] + 2 * lane;
  const int width = w[s];
  int len = 0, start = 0;
  if(r < n) {
    start = ia[r];
    len = ia[r + 1] - start;
  }
  const int self = r < n ? r : 0;
  for(int k = 0; k < width; k++) {
    long d = base + (long)(k >> 1) * (2 * kSlice) + (k & 1);
    double v = k < len ? a[start + k] : 0.;
    if(fill_cols) col[d] = k < len ? ja[start + k] : self;
    val[d] = v;
    if(F32) val32[d] = (float)v;
  }
}

template <int U>
__global__ void __launch_bounds__(256) k_sell_spmv(int n, const long *__restrict__ sp,
                                                   const int *__restrict__ w,
                                                   const int *__restrict__ col,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ y)
{
  int row = blockIdx.x * 256 + threadIdx.x;
  if(row >= n) return;
  double sum;
  sell_row<double, double, U>(row, sp, w, col, val, x, sum);
  y[row] = sum;
}

}  // namespace

int sell_build(const std::vector<int> &h_ia, const DevCSR &A, DevSell &S)
{
  S.release();
  const int n = A.n;
  S.n = n;
  S.nnz = A.nnz;
  S.nslices = (n + kSlice - 1) / kSlice;
  std::vector<long> sp(S.nslices + 1, 0);
  std::vector<int> w(std::max(1, S.nslices), 0);
  for(int s = 0; s < S.nslices; s++) {
    int m = 0;
    for(int r = s * kSlice; r < std::min(n, (s + 1) * kSlice); r++) m = std::max(m, h_ia[r + 1] - h_ia[r]);
    m = (m + 1) & ~1;  // entries go in pairs
    w[s] = m;
    sp[s + 1] = sp[s] + (long)m * kSlice;
  }
  S.nnz_pad = sp[S.nslices];
  GDP_TRY(dalloc(&S.sp, sp.size()));
  GDP_TRY(dalloc(&S.w, w.size()));
  GDP_TRY(dalloc(&S.col, (size_t)std::max(1L, S.nnz_pad)));
  GDP_TRY(dalloc(&S.val, (size_t)std::max(1L, S.nnz_pad)));
  GDP_TRY(h2d(S.sp, sp.data(), sp.size()));
  GDP_TRY(h2d(S.w, w.data(), w.size()));
  if(n) {
    int rows = S.nslices * kSlice;
    hipLaunchKernelGGL(k_sell_fill<false>, dim3((rows + 255) / 256), dim3(256), 0, ctx().stream, n,
                       A.ia, A.ja, A.a, S.sp, S.w, S.col, S.val, (float *)nullptr, true);
  }
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int sell_copy_values(const DevCSR &A, DevSell &S, bool also_f32)
{
  if(!S.n) return GDP_OK;
  if(also_f32 && !S.val32) GDP_TRY(dalloc(&S.val32, (size_t)std::max(1L, S.nnz_pad)));
  int rows = S.nslices * kSlice;
  if(also_f32)
    hipLaunchKernelGGL(k_sell_fill<true>, dim3((rows + 255) / 256), dim3(256), 0, ctx().stream, S.n,
                       A.ia, A.ja, A.a, S.sp, S.w, S.col, S.val, S.val32, false);
  else
    hipLaunchKernelGGL(k_sell_fill<false>, dim3((rows + 255) / 256), dim3(256), 0, ctx().stream, S.n,
                       A.ia, A.ja, A.a, S.sp, S.w, S.col, S.val, (float *)nullptr, false);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int sell_spmv(const DevSell &S, const double *x, double *y)
{
  if(!S.n) return GDP_OK;
  // GDP_SELL_UNROLL=4 keeps 4 pairs per lane in flight instead of 2
  // (launch-shape experiments)
  static int unroll = -1;
  if(unroll < 0) {
    const char *e = std::getenv("GDP_SELL_UNROLL");
    unroll = (e && std::atoi(e) == 4) ? 4 : 2;
  }
  if(unroll == 4)
    hipLaunchKernelGGL(k_sell_spmv<4>, dim3((S.n + 255) / 256), dim3(256), 0, ctx().stream, S.n,
                       S.sp, S.w, S.col, S.val, x, y);
  else
    hipLaunchKernelGGL(k_sell_spmv<2>, dim3((S.n + 255) / 256), dim3(256), 0, ctx().stream, S.n,
                       S.sp, S.w, S.col, S.val, x, y);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

}  // namespace gdp
