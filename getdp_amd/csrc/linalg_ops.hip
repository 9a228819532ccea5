// Whole-vector and whole-matrix arithmetic of the LinAlg interface, as the
// Sparskit backend defines it (LinAlg_SPARSKIT.cpp:66-800 over the helpers of
// Sparskit.cpp:797-1215), on the device:
//   V1 += V2, V1 += d * V2, V1 -= V2 or V2 = V1 - V2, V *= d, swap, set all
//   entries, dot product, 2-norm, inf-norm, zeroing the dummy DOFs;
//   M1 += M2 (or M2 += M1), M1 += d * M2, M *= d.
// Elementwise results are those of the sequential loops bit for bit (one
// rounding per operation, no contraction).  The dot product is a fixed
// two-stage reduction, deterministic but not in the reference's sequential
// order.  The inf-norm, like the reference's `if(fabs(u) > norm)` loop,
// ignores NaN entries.  Argument rules follow the reference: the result must
// alias the operand it names ("Wrong arguments in ..." otherwise).
#include <cmath>
#include <vector>

#include "gdp_device.h"

namespace gdp {

namespace {

constexpr int kB = 256;

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kB - 1) / kB)); }

// op 0: u += v; 1: u += d v; 2: u -= v; 3: v = u - v; 4: u *= d; 5: swap;
// 6: u = d
__global__ void __launch_bounds__(kB) k_vec_op(long n, int op, double d, double *__restrict__ u,
                                               double *__restrict__ v)
{
  const long i = (long)blockIdx.x * kB + threadIdx.x;
  if(i >= n) return;
  switch(op) {
  case 0: u[i] = u[i] + v[i]; break;
  case 1: u[i] = u[i] + d * v[i]; break;
  case 2: u[i] = u[i] - v[i]; break;
  case 3: v[i] = u[i] - v[i]; break;
  case 4: u[i] = u[i] * d; break;
  case 5: {
    const double t = u[i];
    u[i] = v[i];
    v[i] = t;
    break;
  }
  default: u[i] = d; break;
  }
}

__global__ void __launch_bounds__(kB) k_dot_partial(long n, const double *__restrict__ u,
                                                    const double *__restrict__ v, double *part)
{
  double s = 0;
  for(long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) s += u[i] * v[i];
  s = block_sum<double, kB>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kB) k_absmax_partial(long n, const double *__restrict__ u, double *part)
{
  __shared__ double sh[kB];
  double m = 0.;
  for(long i = (long)blockIdx.x * kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) {
    const double a = fabs(u[i]);
    if(a > m) m = a;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for(int o = kB / 2; o > 0; o >>= 1) {
    if((int)threadIdx.x < o && sh[threadIdx.x + o] > sh[threadIdx.x]) sh[threadIdx.x] = sh[threadIdx.x + o];
    __syncthreads();
  }
  if(threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kB) k_zero_masked(long n, const int *__restrict__ mask, double *__restrict__ u)
{
  const long i = (long)blockIdx.x * kB + threadIdx.x;
  if(i < n && mask[i] == 1) u[i] = 0.;
}

int vec_op(int op, double d, gdp_vector *u, gdp_vector *v)
{
  // staged additions count as already made (LinAlg_AddDoubleInVector adds
  // straight into the vector in the reference); u = d drops u's
  if(op == 6) u->staged.clear();
  else GDP_TRY(vec_flush(u));
  if(v) GDP_TRY(vec_flush(v));
  if(u->n)
    hipLaunchKernelGGL(k_vec_op, grid_for(u->n), dim3(kB), 0, ctx().stream, (long)u->n, op, d, u->d,
                       v ? v->d : nullptr);
  GDP_HIP(hipGetLastError());
  return sync();
}

// Every reduction launches kParts blocks (grid-stride over the entries), each
// writing one partial; the host combines the partials in block order.
constexpr int kParts = 512;

// The reduction over the partial buffer `part` (kParts doubles on the device).
int reduce_into(int n, const double *u, const double *v, bool absmax, double *part, double *out)
{
  if(absmax) hipLaunchKernelGGL(k_absmax_partial, dim3(kParts), dim3(kB), 0, ctx().stream, (long)n, u, part);
  else hipLaunchKernelGGL(k_dot_partial, dim3(kParts), dim3(kB), 0, ctx().stream, (long)n, u, v, part);
  GDP_HIP(hipGetLastError());
  double h[kParts];
  GDP_TRY(d2h(h, part, kParts));
  double s = 0.;
  for(double x : h) {
    if(absmax) {
      if(x > s) s = x;
    }
    else
      s += x;
  }
  *out = s;
  return GDP_OK;
}

int reduce(int n, const double *u, const double *v, bool absmax, double *out)
{
  if(!n) {
    *out = 0.;
    return GDP_OK;
  }
  double *part = nullptr;
  GDP_TRY(dalloc(&part, kParts));
  int r = reduce_into(n, u, v, absmax, part, out);
  dfree(part);
  return r;
}

// ---- the allocation-free dot product -----------------------------------------
// gdp_vector_dot_noalloc keeps its kParts partial sums (4 KiB) in this array
// instead of a hipMalloc'd buffer.  Every GPU the library runs on holds its own
// copy until the process ends.  In host code the array's name is not a device
// address: the address hipGetSymbolAddress returns for each device, looked up
// once per device at gdp_init with that device current, is what host code
// passes to the kernels.
static __device__ __align__(256) double scratch[kParts];
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};

__global__ void __launch_bounds__(kB) k_mat_axpy(long nnz, double d, int use_d, double *__restrict__ a,
                                                 const double *__restrict__ b)
{
  const long k = (long)blockIdx.x * kB + threadIdx.x;
  if(k < nnz) a[k] = use_d ? a[k] + d * b[k] : a[k] + b[k];
}

__global__ void __launch_bounds__(kB) k_mat_scale(long nnz, double d, double *__restrict__ a)
{
  const long k = (long)blockIdx.x * kB + threadIdx.x;
  if(k < nnz) a[k] = a[k] * d;
}

// dst += (d *) src: on the device when the patterns agree, else by a merge of
// the two sorted patterns on the host (each entry receives at most one
// addition, so the order of the reference's list insertions does not matter)
int mat_add(gdp_matrix *dst, const gdp_matrix *src, double d, bool use_d)
{
  if(dst->n != src->n) return set_error(GDP_E_ARG, "matrices of different size");
  if(!dst->csr.ia || !src->csr.ia) return set_error(GDP_E_STATE, "matrix not assembled");
  if(dst->h_ia == src->h_ia && dst->h_ja == src->h_ja) {
    if(dst->csr.nnz)
      hipLaunchKernelGGL(k_mat_axpy, grid_for(dst->csr.nnz), dim3(kB), 0, ctx().stream, dst->csr.nnz, d,
                         (int)use_d, dst->csr.a, src->csr.a);
    GDP_HIP(hipGetLastError());
    dst->value_version = next_version();
    return sync();
  }
  const int n = dst->n;
  std::vector<double> av((size_t)dst->csr.nnz), bv((size_t)src->csr.nnz);
  if(!av.empty()) GDP_TRY(d2h(av.data(), dst->csr.a, av.size()));
  if(!bv.empty()) GDP_TRY(d2h(bv.data(), src->csr.a, bv.size()));
  std::vector<int> ia(1, 0), ja;
  std::vector<double> a;
  const std::vector<int> &pa = dst->h_ia, &ca = dst->h_ja, &pb = src->h_ia, &cb = src->h_ja;
  for(int i = 0; i < n; i++) {
    int p = pa[i], q = pb[i];
    while(p < pa[i + 1] || q < pb[i + 1]) {
      const int c1 = p < pa[i + 1] ? ca[p] : n, c2 = q < pb[i + 1] ? cb[q] : n;
      if(c1 == c2) {
        ja.push_back(c1);
        a.push_back(use_d ? av[p] + d * bv[q] : av[p] + bv[q]);
        p++;
        q++;
      }
      else if(c1 < c2) {
        ja.push_back(c1);
        a.push_back(av[p++]);
      }
      else {
        ja.push_back(c2);
        a.push_back(use_d ? d * bv[q] : bv[q]);  // a new entry holds the added value itself
        q++;
      }
    }
    ia.push_back((int)ja.size());
  }
  return gdp_matrix_set_csr(dst, n, ia.data(), ja.data(), a.data());
}

}  // namespace

// Called by gdp_init with `device` current.
int noalloc_setup(int device)
{
  if(device < 0 || device >= kMaxDevices) return set_error(GDP_E_ARG, "device %d: at most %d devices", device, kMaxDevices);
  if(g_scratch[device]) return GDP_OK;
  void *p = nullptr;
  GDP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = static_cast<double *>(p);
  return GDP_OK;
}

}  // namespace gdp

using namespace gdp;

#define GDP_VEC2(v1, v2, what)                                                  \
  do {                                                                          \
    if(!(v1) || !(v2)) return set_error(GDP_E_ARG, what ": null argument");    \
    if((v1)->n != (v2)->n) return set_error(GDP_E_ARG, what ": size mismatch"); \
  } while(0)

extern "C" int gdp_vector_add(gdp_vector *v1, gdp_vector *v2, gdp_vector *v3)
{
  GDP_VEC2(v1, v2, "gdp_vector_add");
  if(v3 != v1) return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_AddVectorVector'");
  return vec_op(0, 0., v1, v2);
}

extern "C" int gdp_vector_add_scaled(gdp_vector *v1, gdp_vector *v2, double d, gdp_vector *v3)
{
  GDP_VEC2(v1, v2, "gdp_vector_add_scaled");
  if(v3 != v1) return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_AddVectorProdVectorDouble'");
  return vec_op(1, d, v1, v2);
}

extern "C" int gdp_vector_sub(gdp_vector *v1, gdp_vector *v2, gdp_vector *v3)
{
  GDP_VEC2(v1, v2, "gdp_vector_sub");
  if(v3 == v1) return vec_op(2, 0., v1, v2);
  if(v3 == v2) return vec_op(3, 0., v1, v2);
  return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_SubVectorVector'");
}

extern "C" int gdp_vector_scale(gdp_vector *v1, double d, gdp_vector *v2)
{
  if(!v1) return set_error(GDP_E_ARG, "gdp_vector_scale: null argument");
  if(v2 != v1) return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_ProdVectorDouble'");
  return vec_op(4, d, v1, nullptr);
}

extern "C" int gdp_vector_swap(gdp_vector *v1, gdp_vector *v2)
{
  if(!v1 || !v2) return set_error(GDP_E_ARG, "gdp_vector_swap: null argument");
  if(v1->n != v2->n) return set_error(GDP_E_ARG, "Cannot swap vectors of different size");
  if(v1 == v2) return GDP_OK;
  return vec_op(5, 0., v1, v2);
}

extern "C" int gdp_vector_set_all(gdp_vector *v, double d)
{
  if(!v) return set_error(GDP_E_ARG, "gdp_vector_set_all: null argument");
  return vec_op(6, d, v, nullptr);
}

extern "C" int gdp_vector_dot(const gdp_vector *v1, const gdp_vector *v2, double *d)
{
  if(!v1 || !v2 || !d) return set_error(GDP_E_ARG, "gdp_vector_dot: null argument");
  if(v1->n != v2->n) return set_error(GDP_E_ARG, "gdp_vector_dot: size mismatch");
  GDP_TRY(vec_flush(v1));
  GDP_TRY(vec_flush(v2));
  return reduce(v1->n, v1->d, v2->d, false, d);
}

extern "C" int gdp_vector_dot_noalloc(const gdp_vector *v1, const gdp_vector *v2, double *d)
{
  if(!v1 || !v2 || !d) return set_error(GDP_E_ARG, "gdp_vector_dot_noalloc: null argument");
  if(v1->n != v2->n) return set_error(GDP_E_ARG, "gdp_vector_dot_noalloc: size mismatch");
  // initialising the library or applying staged additions would allocate
  if(!ctx().ready) return set_error(GDP_E_STATE, "gdp_vector_dot_noalloc: call gdp_init first");
  if(!v1->staged.empty() || !v2->staged.empty())
    return set_error(GDP_E_STATE, "gdp_vector_dot_noalloc: a vector has pending additions (gdp_vector_assemble first)");
  double *part = g_scratch[ctx().device];
  if(!part) return set_error(GDP_E_STATE, "gdp_vector_dot_noalloc: no scratch address for device %d", ctx().device);
  if(!v1->n) {
    *d = 0.;
    return GDP_OK;
  }
  // any length fits: the launch is always kParts blocks, one partial each
  return reduce_into(v1->n, v1->d, v2->d, false, part, d);
}

extern "C" int gdp_vector_norm_inf(const gdp_vector *v, double *norm)
{
  if(!v || !norm) return set_error(GDP_E_ARG, "gdp_vector_norm_inf: null argument");
  GDP_TRY(vec_flush(v));
  return reduce(v->n, v->d, nullptr, true, norm);
}

extern "C" int gdp_vector_dummy(gdp_vector *v, const int *dummy)
{
  if(!v) return set_error(GDP_E_ARG, "gdp_vector_dummy: null argument");
  GDP_TRY(vec_flush(v));
  if(!dummy || !v->n) return GDP_OK;  // no dummy DOFs
  int *dm = nullptr;
  GDP_TRY(dalloc(&dm, (size_t)v->n));
  int r = h2d(dm, dummy, (size_t)v->n);
  if(!r) {
    hipLaunchKernelGGL(k_zero_masked, grid_for(v->n), dim3(kB), 0, ctx().stream, (long)v->n, dm, v->d);
    r = hipGetLastError() == hipSuccess ? sync() : set_error(GDP_E_HIP, "k_zero_masked launch");
  }
  dfree(dm);
  return r;
}

extern "C" int gdp_matrix_size(const gdp_matrix *A, int *n, int *m)
{
  if(!A || !n || !m) return set_error(GDP_E_ARG, "gdp_matrix_size: null argument");
  *n = A->n;
  *m = A->m;
  return GDP_OK;
}

extern "C" int gdp_matrix_add(gdp_matrix *m1, gdp_matrix *m2, gdp_matrix *m3)
{
  if(!m1 || !m2) return set_error(GDP_E_ARG, "gdp_matrix_add: null argument");
  if(m3 == m1) return mat_add(m1, m2, 0., false);
  if(m3 == m2) return mat_add(m2, m1, 0., false);
  return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_AddMatrixMatrix'");
}

extern "C" int gdp_matrix_add_scaled(gdp_matrix *m1, gdp_matrix *m2, double d, gdp_matrix *m3)
{
  if(!m1 || !m2) return set_error(GDP_E_ARG, "gdp_matrix_add_scaled: null argument");
  if(m3 != m1) return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_AddMatrixProdMatrixDouble'");
  return mat_add(m1, m2, d, true);
}

extern "C" int gdp_matrix_scale(gdp_matrix *m1, double d, gdp_matrix *m2)
{
  if(!m1) return set_error(GDP_E_ARG, "gdp_matrix_scale: null argument");
  if(m2 != m1) return set_error(GDP_E_ARG, "Wrong arguments in 'LinAlg_ProdMatrixDouble'");
  if(!m1->csr.ia) return set_error(GDP_E_STATE, "matrix not assembled");
  if(m1->csr.nnz)
    hipLaunchKernelGGL(k_mat_scale, grid_for(m1->csr.nnz), dim3(kB), 0, ctx().stream, m1->csr.nnz, d, m1->csr.a);
  GDP_HIP(hipGetLastError());
  m1->value_version = next_version();
  return sync();
}
