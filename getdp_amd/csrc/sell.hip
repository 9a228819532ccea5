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
__global__ void __launch_bounds__(256) k_sell_spmv(SellOp<double> A, const double *__restrict__ x,
                                                   double *__restrict__ y)
{
  int row = blockIdx.x * 256 + threadIdx.x;
  if(row >= A.n) return;
  double sum;
  sell_row<double, double, U>(row, A, x, sum);
  y[row] = sum;
}

// One wave per slice: for each pair group, the smallest column becomes the
// base and the lanes store their two offsets; *wide is set when a group spans
// 65536 columns or more.
__device__ inline int wave_min_i(int v)
{
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_max_i(int v)
{
#pragma unroll
  for(int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ void __launch_bounds__(64) k_sell_compress(const long *__restrict__ sp,
                                                      const int *__restrict__ w,
                                                      const int *__restrict__ col,
                                                      unsigned *__restrict__ c16,
                                                      int *__restrict__ cbase, int *__restrict__ wide)
{
  const int s = blockIdx.x, lane = threadIdx.x;
  const long base = sp[s];
  const int nq = w[s] >> 1;
  int bad = 0;
  for(int q = 0; q < nq; q++) {
    const long d = base + 128L * q + 2 * lane;
    const int c0 = col[d], c1 = col[d + 1];
    const int lo = wave_min_i(min(c0, c1)), hi = wave_max_i(max(c0, c1));
    if(hi - lo > 65535) bad = 1;
    c16[(base >> 1) + 64L * q + lane] = (unsigned)(c0 - lo) | ((unsigned)(c1 - lo) << 16);
    if(lane == 0) cbase[(base >> 7) + q] = lo;
  }
  if(bad && lane == 0) atomicOr(wide, 1);
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
  // GDP_SELL_COL32=1 keeps the 32-bit columns (A/B measurements)
  const char *e32 = std::getenv("GDP_SELL_COL32");
  if(!n || (e32 && e32[0] && e32[0] != '0')) return GDP_OK;
  GDP_TRY(dalloc(&S.c16, (size_t)std::max(1L, S.nnz_pad / 2)));
  GDP_TRY(dalloc(&S.cbase, (size_t)std::max(1L, S.nnz_pad / 128)));
  int *wide = nullptr;
  GDP_TRY(dalloc(&wide, 1));
  int zero = 0, h_wide = 1;
  GDP_TRY(h2d(wide, &zero, 1));
  hipLaunchKernelGGL(k_sell_compress, dim3(S.nslices), dim3(64), 0, ctx().stream, S.sp, S.w, S.col,
                     S.c16, S.cbase, wide);
  GDP_HIP(hipGetLastError());
  int rc = d2h(&h_wide, wide, 1);
  dfree(wide);
  GDP_TRY(rc);
  if(h_wide) {
    dfree(S.c16);
    dfree(S.cbase);
  }
  else
    dfree(S.col);
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
    hipLaunchKernelGGL(k_sell_spmv<4>, dim3((S.n + 255) / 256), dim3(256), 0, ctx().stream,
                       sell_op<double>(S), x, y);
  else
    hipLaunchKernelGGL(k_sell_spmv<2>, dim3((S.n + 255) / 256), dim3(256), 0, ctx().stream,
                       sell_op<double>(S), x, y);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

}  // namespace gdp
