// Device-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "gdp_internal.h"

#define GDP_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if(e_ != hipSuccess)                                                       \
      return ::gdp::set_error(GDP_E_HIP, "%s failed: %s (%s:%d)", #call,       \
                              hipGetErrorString(e_), __FILE__, __LINE__);      \
  } while(0)

#define GDP_TRY(call)                                                          \
  do {                                                                         \
    int r_ = (call);                                                           \
    if(r_ < 0) return r_;                                                      \
  } while(0)

namespace gdp {

constexpr int kWave = 64;

struct Context {
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 256;
  // small pinned host buffer for control-block readback
  void *pinned = nullptr;
  size_t pinned_bytes = 0;
};
Context &ctx();
int ensure_ready();

// Device allocation helpers (return GDP status).
template <typename T> int dalloc(T **p, size_t n)
{
  *p = nullptr;
  if(n == 0) return GDP_OK;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if(e != hipSuccess)
    return set_error(GDP_E_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T),
                     hipGetErrorString(e));
  return GDP_OK;
}
template <typename T> void dfree(T *&p)
{
  if(p) (void)hipFree((void *)p);
  p = nullptr;
}
template <typename T> int h2d(T *d, const T *h, size_t n)
{
  if(!n) return GDP_OK;
  GDP_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx().stream));
  return GDP_OK;
}
template <typename T> int d2h(T *h, const T *d, size_t n)
{
  if(!n) return GDP_OK;
  GDP_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx().stream));
  GDP_HIP(hipStreamSynchronize(ctx().stream));
  return GDP_OK;
}
int sync();
// A fresh pattern / value version number for gdp_matrix.
uint64_t next_version();

// ---- reductions -----------------------------------------------------------
template <typename T> __device__ inline T wave_sum(T v)
{
#pragma unroll
  for(int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
  return v;
}

// Deterministic block sum (fixed tree order); result valid in thread 0.
template <typename T, int BLOCK> __device__ inline T block_sum(T v)
{
  __shared__ T s[BLOCK / kWave];
  v = wave_sum(v);
  int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if(l == 0) s[w] = v;
  __syncthreads();
  T r = 0;
  if(threadIdx.x == 0) {
#pragma unroll
    for(int i = 0; i < BLOCK / kWave; i++) r += s[i];
  }
  __syncthreads();
  return r;
}

// Sum nparts partials (any block size), deterministic, result in thread 0.
template <typename T, int BLOCK> __device__ inline T sum_partials(const T *p, int nparts)
{
  T v = 0;
  for(int i = threadIdx.x; i < nparts; i += BLOCK) v += p[i];
  return block_sum<T, BLOCK>(v);
}

// ---- objects ---------------------------------------------------------------
struct DevCSR {
  int n = 0;
  long nnz = 0;
  int *ia = nullptr, *ja = nullptr;
  double *a = nullptr;
  void release()
  {
    dfree(ia);
    dfree(ja);
    dfree(a);
    n = 0;
    nnz = 0;
  }
};

// Sliced-ELLPACK copy of a CSR matrix (layout described in sell.h).
struct DevSell {
  int n = 0, nslices = 0;
  long nnz = 0, nnz_pad = 0;
  long *sp = nullptr;     // slice start offsets (nslices + 1)
  int *w = nullptr;       // slice widths
  int *col = nullptr;     // nnz_pad (null when the columns are compressed)
  unsigned *c16 = nullptr;  // compressed columns: nnz_pad / 2 words
  int *cbase = nullptr;     // compressed columns: one base per pair group
  double *val = nullptr;  // nnz_pad
  float *val32 = nullptr;
  void release()
  {
    dfree(sp);
    dfree(w);
    dfree(col);
    dfree(c16);
    dfree(cbase);
    dfree(val);
    dfree(val32);
    n = nslices = 0;
    nnz = nnz_pad = 0;
  }
  ~DevSell() { release(); }
};

// y = A x (FP64 CSR), launched on the library stream.
int spmv(const DevCSR &A, const double *x, double *y);
int spmv_f32(int n, const int *ia, const int *ja, const float *a, const float *x, float *y,
             long nnz);
// Choose lanes-per-row from the average row length.
int spmv_lanes(long nnz, int n);

}  // namespace gdp

struct gdp_vector {
  int n = 0;
  double *d = nullptr;
  // LinAlg_AddDoubleInVector calls not yet applied; every entry point that
  // reads the vector applies them first (vec_flush), every one that
  // overwrites it drops them, so the vector behaves as if each addition
  // had gone straight into it (LinAlg_SPARSKIT.cpp:356-364)
  mutable std::vector<std::pair<int, double>> staged;
};

namespace gdp {
// Apply the vector's staged additions in call order (a no-op without any).
int vec_flush(const gdp_vector *v);
// Per-device set-up of gdp_vector_dot_noalloc's scratch address (gdp_init).
int noalloc_setup(int device);
}  // namespace gdp

struct gdp_matrix {
  int n = 0, m = 0;
  gdp::DevCSR csr;
  gdp::DevSell sell;                    // SpMV layout, rebuilt lazily
  uint64_t sell_pattern = 0, sell_values = 0;
  std::vector<int> h_ia, h_ja;          // host copy of the pattern
  uint64_t pattern_version = 0;         // bumps when the pattern changes
  uint64_t value_version = 0;           // bumps when values change
  struct Staged {
    int i, j;
    double v;
  };
  std::vector<Staged> staged;
};
