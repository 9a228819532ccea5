// Library runtime: device context, errors, vectors, matrices and CSR SpMV.
//
// Replaces the storage half of the reference's Sparskit backend
// (contrib/Sparskit/Sparskit.cpp:662-1100: init/zero/add/prod on per-row
// linked lists) with device-resident CSR: entries added through the LinAlg
// interface are staged on the host and turned into one sorted CSR at
// assemble time; the batched assembly path (fe.hip) fills a fixed pattern
// directly on the device.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <cmath>
#include <numeric>

#include "gdp_device.h"
#include "sell.h"

namespace gdp {

namespace {
thread_local char g_err[1024] = "";
}

int set_error(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

Context &ctx()
{
  static Context c;
  return c;
}

int ensure_ready()
{
  if(!ctx().ready) {
    int r = gdp_init(0);
    if(r) return r;
  }
  return GDP_OK;
}

int sync()
{
  GDP_HIP(hipStreamSynchronize(ctx().stream));
  return GDP_OK;
}

// ---- CSR SpMV ----------------------------------------------------------------
// G lanes of a wave64 cooperate on one row: lane l reads entries
// ia[r]+l, ia[r]+l+G, ... (consecutive lanes read consecutive entries, so a
// group streams its row's values and column indices as one contiguous
// segment); the partial sums are combined with wave shuffles.  Values and
// column indices are read once (non-temporal), x through the caches.
template <int G, typename TV, typename TX>
__global__ void __launch_bounds__(256) k_spmv(int n, const int *__restrict__ ia,
                                              const int *__restrict__ ja,
                                              const TV *__restrict__ a,
                                              const TX *__restrict__ x,
                                              TX *__restrict__ y)
{
  long gt = (long)blockIdx.x * 256 + threadIdx.x;
  int row = (int)(gt / G), lane = (int)(gt % G);
  if(row >= n) return;
  int s = ia[row], e = ia[row + 1];
  TX sum = 0;
  for(int k = s + lane; k < e; k += G) {
    TV v = __builtin_nontemporal_load(a + k);
    int c = __builtin_nontemporal_load(ja + k);
    sum += (TX)v * x[c];
  }
#pragma unroll
  for(int o = G / 2; o > 0; o >>= 1) sum += __shfl_down(sum, o, G);
  if(lane == 0) y[row] = sum;
}

int spmv_lanes(long nnz, int n)
{
  double avg = n ? (double)nnz / n : 1.0;
  if(avg <= 3) return 2;
  if(avg <= 10) return 4;
  if(avg <= 24) return 8;
  if(avg <= 64) return 16;
  return 32;
}

template <typename TV, typename TX>
static int launch_spmv(int n, const int *ia, const int *ja, const TV *a, const TX *x, TX *y,
                       long nnz)
{
  if(n == 0) return GDP_OK;
  int G = spmv_lanes(nnz, n);
  long threads = (long)n * G;
  dim3 grid((unsigned)((threads + 255) / 256)), block(256);
  hipStream_t s = ctx().stream;
  switch(G) {
  case 2: hipLaunchKernelGGL((k_spmv<2, TV, TX>), grid, block, 0, s, n, ia, ja, a, x, y); break;
  case 4: hipLaunchKernelGGL((k_spmv<4, TV, TX>), grid, block, 0, s, n, ia, ja, a, x, y); break;
  case 8: hipLaunchKernelGGL((k_spmv<8, TV, TX>), grid, block, 0, s, n, ia, ja, a, x, y); break;
  case 16: hipLaunchKernelGGL((k_spmv<16, TV, TX>), grid, block, 0, s, n, ia, ja, a, x, y); break;
  default: hipLaunchKernelGGL((k_spmv<32, TV, TX>), grid, block, 0, s, n, ia, ja, a, x, y); break;
  }
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int spmv(const DevCSR &A, const double *x, double *y)
{
  return launch_spmv<double, double>(A.n, A.ia, A.ja, A.a, x, y, A.nnz);
}

int spmv_f32(int n, const int *ia, const int *ja, const float *a, const float *x, float *y,
             long nnz)
{
  return launch_spmv<float, float>(n, ia, ja, a, x, y, nnz);
}

// ---- small vector kernels --------------------------------------------------------
__global__ void k_axpy_scatter(int n, const int *idx, const double *val, double *v)
{
  // staged adds are applied by one thread in call order (duplicates allowed)
  if(blockIdx.x == 0 && threadIdx.x == 0)
    for(int k = 0; k < n; k++) v[idx[k]] += val[k];
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_sumsq_partial(int n, const double *x, double *part)
{
  double s = 0;
  for(long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK)
    s += x[i] * x[i];
  s = block_sum<double, BLOCK>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

}  // namespace gdp

using namespace gdp;

// =============================================================================
// C-ABI: library
// =============================================================================
extern "C" const char *gdp_last_error(void) { return g_err; }

extern "C" int gdp_init(int device)
{
  Context &c = ctx();
  if(c.ready) {
    if(device == c.device) return GDP_OK;
    GDP_TRY(gdp_finalize());
  }
  int ndev = 0;
  GDP_HIP(hipGetDeviceCount(&ndev));
  if(device < 0 || device >= ndev)
    return set_error(GDP_E_ARG, "device %d not available (%d devices)", device, ndev);
  GDP_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  GDP_HIP(hipGetDeviceProperties(&prop, device));
  c.cu_count = prop.multiProcessorCount;
  GDP_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  c.pinned_bytes = 1 << 16;
  GDP_HIP(hipHostMalloc(&c.pinned, c.pinned_bytes, hipHostMallocDefault));
  GDP_TRY(noalloc_setup(device));
  c.device = device;
  c.ready = true;
  return GDP_OK;
}

extern "C" int gdp_finalize(void)
{
  Context &c = ctx();
  if(!c.ready) return GDP_OK;
  (void)hipStreamSynchronize(c.stream);
  (void)hipStreamDestroy(c.stream);
  (void)hipHostFree(c.pinned);
  c = Context();
  return GDP_OK;
}

extern "C" int gdp_device_synchronize(void)
{
  GDP_TRY(ensure_ready());
  return sync();
}

// =============================================================================
// C-ABI: vectors
// =============================================================================
extern "C" int gdp_vector_create(int n, gdp_vector **v)
{
  if(n < 0 || !v) return set_error(GDP_E_ARG, "gdp_vector_create: bad arguments");
  GDP_TRY(ensure_ready());
  auto *p = new gdp_vector;
  p->n = n;
  int r = dalloc(&p->d, (size_t)n);
  if(r) {
    delete p;
    return r;
  }
  if(n) {
    hipError_t e = hipMemsetAsync(p->d, 0, sizeof(double) * n, ctx().stream);
    if(e != hipSuccess) {
      dfree(p->d);
      delete p;
      return set_error(GDP_E_HIP, "hipMemsetAsync: %s", hipGetErrorString(e));
    }
  }
  *v = p;
  return sync();
}

extern "C" int gdp_vector_destroy(gdp_vector *v)
{
  if(!v) return GDP_OK;
  (void)sync();
  dfree(v->d);
  delete v;
  return GDP_OK;
}

extern "C" int gdp_vector_zero(gdp_vector *v)
{
  if(!v) return set_error(GDP_E_ARG, "null vector");
  v->staged.clear();
  if(v->n) GDP_HIP(hipMemsetAsync(v->d, 0, sizeof(double) * v->n, ctx().stream));
  return sync();
}

extern "C" int gdp_vector_size(const gdp_vector *v, int *n)
{
  if(!v || !n) return set_error(GDP_E_ARG, "null argument");
  *n = v->n;
  return GDP_OK;
}

extern "C" int gdp_vector_set(gdp_vector *v, const double *host, int n)
{
  if(!v || (!host && n) || n != v->n) return set_error(GDP_E_ARG, "gdp_vector_set: size %d != %d", n, v ? v->n : -1);
  v->staged.clear();  // setting every entry overwrites earlier additions
  GDP_TRY(h2d(v->d, host, (size_t)n));
  return sync();
}

extern "C" int gdp_vector_get(const gdp_vector *v, double *host, int n)
{
  if(!v || (!host && n) || n != v->n) return set_error(GDP_E_ARG, "gdp_vector_get: size %d != %d", n, v ? v->n : -1);
  GDP_TRY(vec_flush(v));
  return d2h(host, v->d, (size_t)n);
}

extern "C" int gdp_vector_add_double(gdp_vector *v, int i, double d)
{
  if(!v || i < 0 || i >= v->n) return set_error(GDP_E_ARG, "gdp_vector_add_double: index %d", i);
  v->staged.emplace_back(i, d);
  return GDP_OK;
}

int gdp::vec_flush(const gdp_vector *v)
{
  if(!v || v->staged.empty()) return GDP_OK;
  size_t m = v->staged.size();
  std::vector<int> idx(m);
  std::vector<double> val(m);
  for(size_t k = 0; k < m; k++) {
    idx[k] = v->staged[k].first;
    val[k] = v->staged[k].second;
  }
  int *didx = nullptr;
  double *dval = nullptr;
  GDP_TRY(dalloc(&didx, m));
  GDP_TRY(dalloc(&dval, m));
  int r = h2d(didx, idx.data(), m);
  if(!r) r = h2d(dval, val.data(), m);
  if(!r) {
    hipLaunchKernelGGL(k_axpy_scatter, dim3(1), dim3(64), 0, ctx().stream, (int)m, didx, dval, v->d);
    r = sync();
  }
  dfree(didx);
  dfree(dval);
  v->staged.clear();
  return r;
}

extern "C" int gdp_vector_assemble(gdp_vector *v)
{
  if(!v) return set_error(GDP_E_ARG, "null vector");
  return vec_flush(v);
}

extern "C" int gdp_vector_norm2(const gdp_vector *v, double *norm)
{
  if(!v || !norm) return set_error(GDP_E_ARG, "null argument");
  GDP_TRY(vec_flush(v));
  if(v->n == 0) {
    *norm = 0;
    return GDP_OK;
  }
  const int nb = 512;
  double *part = nullptr;
  GDP_TRY(dalloc(&part, nb));
  hipLaunchKernelGGL(k_sumsq_partial<256>, dim3(nb), dim3(256), 0, ctx().stream, v->n, v->d, part);
  std::vector<double> h(nb);
  int r = d2h(h.data(), part, nb);
  dfree(part);
  if(r) return r;
  double s = 0;
  for(double x : h) s += x;
  *norm = std::sqrt(s);
  return GDP_OK;
}

extern "C" int gdp_vector_copy(const gdp_vector *src, gdp_vector *dst)
{
  if(!src || !dst || src->n != dst->n) return set_error(GDP_E_ARG, "gdp_vector_copy: size mismatch");
  GDP_TRY(vec_flush(src));
  if(src != dst) dst->staged.clear();
  if(src->n)
    GDP_HIP(hipMemcpyAsync(dst->d, src->d, sizeof(double) * src->n, hipMemcpyDeviceToDevice,
                           ctx().stream));
  return sync();
}

extern "C" int gdp_vector_device_ptr(gdp_vector *v, double **dptr)
{
  if(!v || !dptr) return set_error(GDP_E_ARG, "null argument");
  *dptr = v->d;
  return GDP_OK;
}

// =============================================================================
// C-ABI: matrices
// =============================================================================
uint64_t gdp::next_version()
{
  static uint64_t v = 0;
  return ++v;
}

extern "C" int gdp_matrix_create(int n, int m, gdp_matrix **A)
{
  if(n < 0 || m < 0 || !A) return set_error(GDP_E_ARG, "gdp_matrix_create: bad arguments");
  if(n != m) return set_error(GDP_E_UNSUPPORTED, "only square matrices (%d x %d)", n, m);
  GDP_TRY(ensure_ready());
  auto *p = new gdp_matrix;
  p->n = n;
  p->m = m;
  *A = p;
  return GDP_OK;
}

extern "C" int gdp_matrix_destroy(gdp_matrix *A)
{
  if(!A) return GDP_OK;
  (void)sync();
  A->csr.release();
  delete A;
  return GDP_OK;
}

extern "C" int gdp_matrix_zero(gdp_matrix *A)
{
  if(!A) return set_error(GDP_E_ARG, "null matrix");
  A->staged.clear();
  if(A->csr.nnz) GDP_HIP(hipMemsetAsync(A->csr.a, 0, sizeof(double) * A->csr.nnz, ctx().stream));
  A->value_version = next_version();
  return sync();
}

extern "C" int gdp_matrix_add_double(gdp_matrix *A, int i, int j, double d)
{
  if(!A || i < 0 || j < 0 || i >= A->n || j >= A->m)
    return set_error(GDP_E_ARG, "gdp_matrix_add_double: index (%d,%d)", i, j);
  A->staged.push_back({i, j, d});
  return GDP_OK;
}

extern "C" int gdp_matrix_add_coo(gdp_matrix *A, long n, const int *i, const int *j,
                                  const double *d)
{
  if(!A || n < 0 || (n && (!i || !j || !d))) return set_error(GDP_E_ARG, "gdp_matrix_add_coo");
  A->staged.reserve(A->staged.size() + (size_t)n);
  for(long k = 0; k < n; k++) {
    if(i[k] < 0 || j[k] < 0 || i[k] >= A->n || j[k] >= A->m)
      return set_error(GDP_E_ARG, "gdp_matrix_add_coo: index (%d,%d)", i[k], j[k]);
    A->staged.push_back({i[k], j[k], d[k]});
  }
  return GDP_OK;
}

static int upload_csr(gdp_matrix *A, const std::vector<int> &ia, const std::vector<int> &ja,
                      const std::vector<double> &a)
{
  long nnz = ia[A->n];
  bool same = (A->csr.nnz == nnz && A->h_ia == ia && A->h_ja == ja);
  if(!same) {
    A->csr.release();
    A->csr.n = A->n;
    A->csr.nnz = nnz;
    GDP_TRY(dalloc(&A->csr.ia, (size_t)A->n + 1));
    GDP_TRY(dalloc(&A->csr.ja, (size_t)nnz));
    GDP_TRY(dalloc(&A->csr.a, (size_t)nnz));
    GDP_TRY(h2d(A->csr.ia, ia.data(), (size_t)A->n + 1));
    GDP_TRY(h2d(A->csr.ja, ja.data(), (size_t)nnz));
    A->h_ia = ia;
    A->h_ja = ja;
    A->pattern_version = next_version();
  }
  GDP_TRY(h2d(A->csr.a, a.data(), (size_t)nnz));
  A->value_version = next_version();
  return sync();
}

// csr_format (Sparskit.cpp:2186-2217): rows = first index, columns sorted,
// duplicate (i, j) entries summed in the order they were added
// (add_matrix_double, Sparskit.cpp:809-862).
extern "C" int gdp_matrix_assemble(gdp_matrix *A)
{
  if(!A) return set_error(GDP_E_ARG, "null matrix");
  if(A->staged.empty()) return GDP_OK;
  const auto &S = A->staged;
  std::vector<size_t> ord(S.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
    return S[x].i != S[y].i ? S[x].i < S[y].i : S[x].j < S[y].j;
  });
  std::vector<int> ia(A->n + 1, 0), ja;
  std::vector<double> a;
  ja.reserve(S.size());
  a.reserve(S.size());
  for(size_t p = 0; p < ord.size();) {
    const auto &e = S[ord[p]];
    double v = e.v;
    size_t q = p + 1;
    while(q < ord.size() && S[ord[q]].i == e.i && S[ord[q]].j == e.j) v += S[ord[q++]].v;
    ja.push_back(e.j);
    a.push_back(v);
    ia[e.i + 1]++;
    p = q;
  }
  for(int r = 0; r < A->n; r++) ia[r + 1] += ia[r];
  A->staged.clear();
  return upload_csr(A, ia, ja, a);
}

extern "C" int gdp_matrix_set_csr(gdp_matrix *A, int n, const int *ia, const int *ja,
                                  const double *a)
{
  if(!A || n != A->n || !ia || (!ja && ia[n]) || (!a && ia[n]))
    return set_error(GDP_E_ARG, "gdp_matrix_set_csr: bad arguments");
  std::vector<int> hia(ia, ia + n + 1), hja(ja, ja + ia[n]);
  std::vector<double> ha(a, a + ia[n]);
  for(int r = 0; r < n; r++)
    for(int k = ia[r]; k < ia[r + 1]; k++)
      if(ja[k] < 0 || ja[k] >= n || (k > ia[r] && ja[k] <= ja[k - 1]))
        return set_error(GDP_E_ARG, "gdp_matrix_set_csr: row %d columns not sorted/unique", r);
  return upload_csr(A, hia, hja, ha);
}

extern "C" int gdp_matrix_nnz(const gdp_matrix *A, long *nnz)
{
  if(!A || !nnz) return set_error(GDP_E_ARG, "null argument");
  *nnz = A->csr.nnz;
  return GDP_OK;
}

extern "C" int gdp_matrix_get_csr(const gdp_matrix *A, int *ia, int *ja, double *a)
{
  if(!A) return set_error(GDP_E_ARG, "null matrix");
  if(ia) std::copy(A->h_ia.begin(), A->h_ia.end(), ia);
  if(ja) std::copy(A->h_ja.begin(), A->h_ja.end(), ja);
  if(a) return d2h(a, A->csr.a, (size_t)A->csr.nnz);
  return GDP_OK;
}

// (Re)build the matrix's SELL copy when its pattern or values changed.
static int matrix_sell(gdp_matrix *A)
{
  if(A->sell_pattern != A->pattern_version || !A->sell.n) {
    GDP_TRY(sell_build(A->h_ia, A->csr, A->sell));
    A->sell_pattern = A->pattern_version;
    A->sell_values = A->value_version;
  }
  else if(A->sell_values != A->value_version) {
    GDP_TRY(sell_copy_values(A->csr, A->sell, false));
    A->sell_values = A->value_version;
  }
  return GDP_OK;
}

extern "C" int gdp_matrix_sell_info(gdp_matrix *A, long *nnz_pad, int *compressed)
{
  if(!A || !nnz_pad || !compressed) return set_error(GDP_E_ARG, "gdp_matrix_sell_info: null argument");
  if(!A->csr.ia) return set_error(GDP_E_STATE, "matrix has no assembled pattern");
  GDP_TRY(matrix_sell(A));
  *nnz_pad = A->sell.nnz_pad;
  *compressed = A->sell.c16 != nullptr;
  return sync();
}

extern "C" int gdp_prod_matrix_vector(gdp_matrix *A, const gdp_vector *x, gdp_vector *y)
{
  if(!A || !x || !y || x->n != A->m || y->n != A->n)
    return set_error(GDP_E_ARG, "gdp_prod_matrix_vector: size mismatch");
  if(!A->csr.ia) return set_error(GDP_E_STATE, "matrix has no assembled pattern");
  GDP_TRY(vec_flush(x));
  if(y != x) y->staged.clear();  // y is overwritten
  GDP_TRY(matrix_sell(A));
  GDP_TRY(sell_spmv(A->sell, x->d, y->d));
  return sync();
}

extern "C" int gdp_bench_spmv(gdp_matrix *A, const gdp_vector *x, gdp_vector *y, int reps,
                              double *avg_ms)
{
  if(!A || !x || !y || reps < 1 || !avg_ms || x->n != A->m || y->n != A->n)
    return set_error(GDP_E_ARG, "gdp_bench_spmv: bad arguments");
  if(!A->csr.ia) return set_error(GDP_E_STATE, "matrix has no assembled pattern");
  GDP_TRY(vec_flush(x));
  if(y != x) y->staged.clear();
  hipEvent_t e0, e1;
  GDP_HIP(hipEventCreate(&e0));
  GDP_HIP(hipEventCreate(&e1));
  GDP_TRY(matrix_sell(A));
  GDP_TRY(sync());
  GDP_HIP(hipEventRecord(e0, ctx().stream));
  for(int r = 0; r < reps; r++) GDP_TRY(sell_spmv(A->sell, x->d, y->d));
  GDP_HIP(hipEventRecord(e1, ctx().stream));
  GDP_HIP(hipEventSynchronize(e1));
  float ms = 0;
  GDP_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / reps;
  return GDP_OK;
}
