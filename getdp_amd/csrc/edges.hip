// Edge numbering on the device (Geo_CreateEntitiesOfElement,
// GeoEntity.cpp:437-490): the key of an edge is its sorted pair of global node
// numbers, an edge takes the next number at its first encounter in visiting
// order (element after element, local edge after local edge), and an element
// stores it with sign -1 when its local edge runs from the higher node to the
// lower.  A repeated visit of an element is skipped.
//
// First-encounter numbering without a sequential walk:
//   1. every (visit position v, local edge k) occurrence gets index o = 6 v + k
//      and its key; occurrences of repeated visits and past the element's
//      edges get an empty key;
//   2. a stable radix sort of (key, o) puts each edge's occurrences together,
//      first occurrence first;
//   3. a flag per occurrence index marks first occurrences; its exclusive scan
//      over o (the visiting order) is the edge number minus one;
//   4. each occurrence finds its run's head (max-scan of head positions) and
//      writes the signed number; run heads write the edge's node pair.
// The numbers and signs are those of the sequential walk, by construction.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdlib>

#include "gdp_device.h"

namespace gdp {

namespace {

// Local edges of Den_Triangle / Den_Tetrahedron (GeoEntity.h:93-94, 214-215)
// and of a line, as in the host walk.
__constant__ int cTriEdge[3][2] = {{0, 1}, {0, 2}, {1, 2}};
__constant__ int cTetEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int cLineEdge[1][2] = {{0, 1}};
constexpr unsigned long long kNoKey = ~0ull;

__device__ inline int local_edges(int et, const int (**tab)[2])
{
  if(et == GDP_TET) {
    *tab = cTetEdge;
    return 6;
  }
  if(et == GDP_TRI) {
    *tab = cTriEdge;
    return 3;
  }
  if(et == GDP_LINE) {
    *tab = cLineEdge;
    return 1;
  }
  return -1;
}

__global__ void __launch_bounds__(256) k_fill_int(long n, int *p, int v)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i < n) p[i] = v;
}

// firstpos[e] = first visit position of element e; err[0] = first bad position
__global__ void __launch_bounds__(256) k_first_visit(int nvisit, const int *visit, int nelem,
                                                     const int *etype, int *firstpos, int *err)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  if(v >= nvisit) return;
  const int e = visit[v];
  if(e < 0 || e >= nelem) {
    atomicMin(err, v);
    return;
  }
  const int(*tab)[2];
  if(local_edges(etype[e], &tab) < 0) atomicMin(err + 1, v);
  atomicMin(firstpos + e, v);
}

__global__ void __launch_bounds__(256) k_edge_keys(int nvisit, const int *visit, const int *etype,
                                                   const int *conn, const int *firstpos,
                                                   unsigned long long *key, int *val)
{
  const int v = blockIdx.x * 256 + threadIdx.x;
  if(v >= nvisit) return;
  const int e = visit[v];
  const int(*tab)[2] = nullptr;
  const int ne = firstpos[e] == v ? local_edges(etype[e], &tab) : 0;
  for(int k = 0; k < 6; k++) {
    const long o = (long)v * 6 + k;
    unsigned long long kk = kNoKey;
    if(k < ne) {
      const int na = conn[(long)e * 4 + tab[k][0]], nb = conn[(long)e * 4 + tab[k][1]];
      const unsigned lo = (unsigned)(na < nb ? na : nb), hi = (unsigned)(na < nb ? nb : na);
      kk = ((unsigned long long)lo << 32) | hi;
    }
    key[o] = kk;
    val[o] = (int)o;
  }
}

__global__ void __launch_bounds__(256) k_edge_heads(long m, const unsigned long long *key, const int *val,
                                                    int *flag, int *headpos)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i >= m) return;
  const bool head = key[i] != kNoKey && (i == 0 || key[i - 1] != key[i]);
  flag[val[i]] = head ? 1 : 0;
  headpos[i] = head ? (int)i : 0;
}

__global__ void __launch_bounds__(256) k_edge_assign(long m, const unsigned long long *key, const int *val,
                                                     const int *runhead, const int *rank, const int *visit,
                                                     const int *etype, const int *conn, int *edges_out,
                                                     int *edge_nodes)
{
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i >= m || key[i] == kNoKey) return;
  const int o = val[i], v = o / 6, k = o % 6, e = visit[v];
  const int h = runhead[i];
  const int num = rank[val[h]] + 1;
  const int(*tab)[2];
  local_edges(etype[e], &tab);
  const int na = conn[(long)e * 4 + tab[k][0]], nb = conn[(long)e * 4 + tab[k][1]];
  edges_out[(long)e * 6 + k] = na < nb ? num : -num;
  if(h == i && edge_nodes) {
    edge_nodes[2 * (long)(num - 1)] = (int)(key[i] >> 32);
    edge_nodes[2 * (long)(num - 1) + 1] = (int)(key[i] & 0xffffffffu);
  }
}

struct MaxOp {
  __device__ int operator()(int a, int b) const { return a > b ? a : b; }
};

}  // namespace

bool edges_on_device()
{
  const char *e = std::getenv("GDP_HOST_PREP");
  return ctx().ready && !(e && e[0] && e[0] != '0');
}

int number_edges_device(int nelem, const int *etype, const int *conn, int nvisit, const int *visit,
                        int *edges_out, int *edge_nodes, int max_edges)
{
  GDP_TRY(ensure_ready());
  hipStream_t s = ctx().stream;
  const long m = (long)nvisit * 6;
  if(m >= INT_MAX) return set_error(GDP_E_ARG, "gdp_number_edges: %d visits", nvisit);
  int *d_et = nullptr, *d_conn = nullptr, *d_visit = nullptr, *d_first = nullptr, *d_err = nullptr;
  int *d_val = nullptr, *d_val2 = nullptr, *d_flag = nullptr, *d_rank = nullptr, *d_hp = nullptr,
      *d_rh = nullptr, *d_out = nullptr, *d_en = nullptr;
  unsigned long long *d_key = nullptr, *d_key2 = nullptr;
  void *tmp = nullptr;
  int count = 0;
  auto body = [&]() -> int {
    const size_t M = (size_t)std::max(1L, m);
    GDP_TRY(dalloc(&d_et, (size_t)std::max(1, nelem)));
    GDP_TRY(dalloc(&d_conn, (size_t)std::max(1, nelem) * 4));
    GDP_TRY(dalloc(&d_visit, (size_t)std::max(1, nvisit)));
    GDP_TRY(dalloc(&d_first, (size_t)std::max(1, nelem)));
    GDP_TRY(dalloc(&d_err, 2));
    GDP_TRY(dalloc(&d_out, (size_t)std::max(1, nelem) * 6));
    if(nelem) {
      GDP_TRY(h2d(d_et, etype, (size_t)nelem));
      GDP_TRY(h2d(d_conn, conn, (size_t)nelem * 4));
    }
    if(nvisit) GDP_TRY(h2d(d_visit, visit, (size_t)nvisit));
    const int big[2] = {INT_MAX, INT_MAX};
    GDP_TRY(h2d(d_err, big, 2));
    GDP_HIP(hipMemsetAsync(d_out, 0, sizeof(int) * (size_t)std::max(1, nelem) * 6, s));
    if(nelem) hipLaunchKernelGGL(k_fill_int, dim3((nelem + 255) / 256), dim3(256), 0, s, (long)nelem, d_first, INT_MAX);
    if(nvisit)
      hipLaunchKernelGGL(k_first_visit, dim3((nvisit + 255) / 256), dim3(256), 0, s, nvisit, d_visit, nelem, d_et,
                         d_first, d_err);
    int herr[2];
    GDP_TRY(d2h(herr, d_err, 2));
    if(herr[0] != INT_MAX) return set_error(GDP_E_ARG, "visit[%d] out of range", herr[0]);
    if(herr[1] != INT_MAX) return set_error(GDP_E_UNSUPPORTED, "element type %d", etype[visit[herr[1]]]);
    if(!m) return GDP_OK;
    GDP_TRY(dalloc(&d_key, M));
    GDP_TRY(dalloc(&d_key2, M));
    GDP_TRY(dalloc(&d_val, M));
    GDP_TRY(dalloc(&d_val2, M));
    GDP_TRY(dalloc(&d_flag, M));
    GDP_TRY(dalloc(&d_rank, M));
    GDP_TRY(dalloc(&d_hp, M));
    GDP_TRY(dalloc(&d_rh, M));
    hipLaunchKernelGGL(k_edge_keys, dim3((nvisit + 255) / 256), dim3(256), 0, s, nvisit, d_visit, d_et, d_conn,
                       d_first, d_key, d_val);
    // stable sort: the occurrences of an edge keep their visiting order
    size_t b1 = 0, b2 = 0, b3 = 0;
    GDP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, d_key, d_key2, d_val, d_val2, (int)m, 0, 64, s));
    GDP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, d_flag, d_rank, (int)m, s));
    GDP_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, b3, d_hp, d_rh, MaxOp(), (int)m, s));
    size_t bytes = std::max(b1, std::max(b2, b3));
    GDP_HIP(hipMalloc(&tmp, std::max<size_t>(bytes, 16)));
    GDP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, b1, d_key, d_key2, d_val, d_val2, (int)m, 0, 64, s));
    const dim3 gm((unsigned)((m + 255) / 256));
    hipLaunchKernelGGL(k_edge_heads, gm, dim3(256), 0, s, m, d_key2, d_val2, d_flag, d_hp);
    GDP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, b2, d_flag, d_rank, (int)m, s));
    GDP_HIP(hipcub::DeviceScan::InclusiveScan(tmp, b3, d_hp, d_rh, MaxOp(), (int)m, s));
    int last[2];
    GDP_TRY(d2h(last, d_rank + (m - 1), 1));
    GDP_TRY(d2h(last + 1, d_flag + (m - 1), 1));
    count = last[0] + last[1];
    if(count > max_edges) return set_error(GDP_E_ARG, "more than %d edges", max_edges);
    if(edge_nodes && count) GDP_TRY(dalloc(&d_en, (size_t)count * 2));
    hipLaunchKernelGGL(k_edge_assign, gm, dim3(256), 0, s, m, d_key2, d_val2, d_rh, d_rank, d_visit, d_et, d_conn,
                       d_out, d_en);
    GDP_HIP(hipGetLastError());
    if(edge_nodes && count) GDP_TRY(d2h(edge_nodes, d_en, (size_t)count * 2));
    return GDP_OK;
  };
  int rc = body();
  if(!rc && nelem) rc = d2h(edges_out, d_out, (size_t)nelem * 6);
  if(tmp) (void)hipFree(tmp);
  dfree(d_et);
  dfree(d_conn);
  dfree(d_visit);
  dfree(d_first);
  dfree(d_err);
  dfree(d_val);
  dfree(d_val2);
  dfree(d_flag);
  dfree(d_rank);
  dfree(d_hp);
  dfree(d_rh);
  dfree(d_out);
  dfree(d_en);
  dfree(d_key);
  dfree(d_key2);
  return rc ? rc : count;
}

}  // namespace gdp
