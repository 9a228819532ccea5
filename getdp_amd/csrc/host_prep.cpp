// Host-side pre-processing: edge numbering, RCMK renumbering, dependency
// levels, colouring and symmetric permutation of CSR patterns.
//
// These are integer/graph algorithms run once per mesh or per matrix pattern;
// they fix the numbering contracts that the device kernels rely on.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "gdp_internal.h"
#include "host_par.h"

namespace gdp {

namespace {

// Local edges of the reference element tables Den_Triangle / Den_Tetrahedron
// (GeoEntity.h:93-94, 214-215) and of a line.
const int kTriEdge[3][2] = {{0, 1}, {0, 2}, {1, 2}};
const int kTetEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
const int kLineEdge[1][2] = {{0, 1}};

inline uint64_t mix64(uint64_t k)
{
  k ^= k >> 31;
  k *= 0x7fb5d329728ea185ULL;
  k ^= k >> 27;
  k *= 0x81dadef4bc2dd44dULL;
  k ^= k >> 33;
  return k;
}

}  // namespace

// Edge numbering contract of Geo_CreateEntitiesOfElement (GeoEntity.cpp:437-490):
// the key of an edge is its sorted pair of global node numbers; an edge that is
// not yet known gets the next number (first encounter in visiting order); the
// element stores it with sign -1 when the lower node number is the edge's
// second local node.
int number_edges(int nelem, const int *etype, const int *conn, int nvisit,
                 const int *visit, int *edges_out, int *edge_nodes, int max_edges)
{
  size_t cap = 64;
  while(cap < 2 * (size_t)max_edges + 64) cap <<= 1;
  std::vector<uint64_t> keys(cap, 0);
  std::vector<int> nums(cap, 0);
  std::memset(edges_out, 0, sizeof(int) * (size_t)nelem * 6);
  std::vector<char> done((size_t)nelem, 0);
  int count = 0;
  for(int v = 0; v < nvisit; v++) {
    int e = visit[v];
    if(e < 0 || e >= nelem) return set_error(GDP_E_ARG, "visit[%d]=%d out of range", v, e);
    if(done[e]) continue;
    done[e] = 1;
    const int(*tab)[2];
    int ne;
    switch(etype[e]) {
    case GDP_TET: tab = kTetEdge; ne = 6; break;
    case GDP_TRI: tab = kTriEdge; ne = 3; break;
    case GDP_LINE: tab = kLineEdge; ne = 1; break;
    default: return set_error(GDP_E_UNSUPPORTED, "element type %d", etype[e]);
    }
    const int *c = conn + (size_t)e * 4;
    for(int k = 0; k < ne; k++) {
      int na = c[tab[k][0]], nb = c[tab[k][1]];
      uint32_t lo = (uint32_t)std::min(na, nb), hi = (uint32_t)std::max(na, nb);
      uint64_t key = ((uint64_t)lo << 32) | hi;
      size_t h = mix64(key) & (cap - 1);
      while(nums[h] && keys[h] != key) h = (h + 1) & (cap - 1);
      if(!nums[h]) {
        if(count >= max_edges) return set_error(GDP_E_ARG, "more than %d edges", max_edges);
        keys[h] = key;
        nums[h] = ++count;
        if(edge_nodes) {
          edge_nodes[2 * (size_t)(count - 1)] = (int)lo;
          edge_nodes[2 * (size_t)(count - 1) + 1] = (int)hi;
        }
      }
      edges_out[(size_t)e * 6 + k] = (na < nb) ? nums[h] : -nums[h];
    }
  }
  return count;
}

// Reverse Cuthill-McKee as GetDP's Sparskit path computes it
// (cmkreord.f:3-62): starting from node 1, repeat a breadth-first level
// traversal from the minimum-degree node of the last level while the number
// of levels grows (perphn, reordering.f:232-271); unvisited components are
// entered in natural order (BFS with iperm(1) = 0, reordering.f:15-156);
// the final traversal is reversed (rversp) and inverted (exchange).
int rcmk(int n, const int *ia, const int *ja, std::vector<int> &perm,
         std::vector<int> &rperm)
{
  perm.assign(n, 0);
  rperm.assign(n, 0);
  if(n == 0) return 0;
  std::vector<char> seen(n, 0);
  std::vector<int> lev_start;
  auto traverse = [&](int start) {
    std::fill(seen.begin(), seen.end(), 0);
    lev_start.clear();
    int head = 0, tail = 0, scan = 0;
    rperm[tail++] = start;
    seen[start] = 1;
    for(;;) {
      lev_start.push_back(head);
      int level_end = tail;
      for(int p = head; p < level_end; p++) {
        int i = rperm[p];
        for(int k = ia[i]; k < ia[i + 1]; k++) {
          int j = ja[k];
          if(!seen[j]) {
            seen[j] = 1;
            rperm[tail++] = j;
          }
        }
      }
      head = level_end;
      if(head < tail) continue;
      while(scan < n && seen[scan]) scan++;
      if(scan >= n) break;
      head = tail;
      rperm[tail++] = scan;
      seen[scan] = 1;
    }
    lev_start.push_back(tail);
    return (int)lev_start.size() - 1;
  };
  int start = 0, best = 0, nlev = 0;
  for(;;) {
    nlev = traverse(start);
    if(nlev <= best) break;
    int mindeg = lev_start[nlev];
    for(int p = lev_start[nlev - 1]; p < lev_start[nlev]; p++) {
      int node = rperm[p];
      int deg = ia[node + 1] - ia[node];
      if(deg < mindeg) {
        start = node;
        mindeg = deg;
      }
    }
    best = nlev;
  }
  std::reverse(rperm.begin(), rperm.end());
  for(int i = 0; i < n; i++) perm[rperm[i]] = i;
  return nlev;
}

int lower_levels(int n, const int *ia, const int *ja, std::vector<int> &lev)
{
  lev.assign(n, 0);
  int depth = 0;
  for(int i = 0; i < n; i++) {
    int l = 0;
    for(int k = ia[i]; k < ia[i + 1] && ja[k] < i; k++) l = std::max(l, lev[ja[k]] + 1);
    lev[i] = l;
    depth = std::max(depth, l + 1);
  }
  return depth;
}

int upper_levels(int n, const int *ia, const int *ja, std::vector<int> &lev)
{
  lev.assign(n, 0);
  int depth = 0;
  for(int i = n - 1; i >= 0; i--) {
    int l = 0;
    for(int k = ia[i + 1] - 1; k >= ia[i] && ja[k] > i; k--) l = std::max(l, lev[ja[k]] + 1);
    lev[i] = l;
    depth = std::max(depth, l + 1);
  }
  return depth;
}

int greedy_colouring(int n, const int *ia, const int *ja, std::vector<int> &col)
{
  col.assign(n, -1);
  std::vector<int> mark(64, -1);
  int ncol = 0;
  for(int i = 0; i < n; i++) {
    for(int k = ia[i]; k < ia[i + 1]; k++) {
      int c = col[ja[k]];
      if(c >= 0) {
        if(c >= (int)mark.size()) mark.resize(2 * c + 2, -1);
        mark[c] = i;
      }
    }
    int c = 0;
    while(c < (int)mark.size() && mark[c] == i) c++;
    if(c >= (int)mark.size()) mark.resize(2 * c + 2, -1);
    col[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void permute_pattern(int n, const int *ia, const int *ja, const int *perm,
                     std::vector<int> &ib, std::vector<int> &jb,
                     std::vector<int> &map)
{
  ib.assign(n + 1, 0);
  for(int i = 0; i < n; i++) ib[perm[i] + 1] = ia[i + 1] - ia[i];
  for(int i = 0; i < n; i++) ib[i + 1] += ib[i];
  size_t nnz = (size_t)ia[n];
  jb.resize(nnz);
  map.resize(nnz);
  // rows are independent (each fills its own range of jb and map): over
  // threads, each row's (new column, old position) pairs sorted by column
  // (insertion sort: rows are short)
  int *jbp = jb.data(), *mp = map.data();
  const int *ibp = ib.data();
  parallel_ranges((size_t)n, [=](size_t a, size_t b) {
    std::vector<std::pair<int, int>> tmp;
    for(size_t i = a; i < b; i++) {
      tmp.clear();
      for(int k = ia[i]; k < ia[i + 1]; k++) {
        const std::pair<int, int> t(perm[ja[k]], k);
        size_t j = tmp.size();
        tmp.push_back(t);
        while(j > 0 && t < tmp[j - 1]) {
          tmp[j] = tmp[j - 1];
          j--;
        }
        tmp[j] = t;
      }
      int o = ibp[perm[i]];
      for(const auto &t : tmp) {
        jbp[o] = t.first;
        mp[o] = t.second;
        o++;
      }
    }
  });
}

}  // namespace gdp

extern "C" int gdp_number_edges(int nelem, const int *etype, const int *conn,
                                int nvisit, const int *visit, int *edges_out,
                                int *edge_nodes, int max_edges)
{
  if(nelem < 0 || !etype || !conn || !visit || !edges_out)
    return gdp::set_error(GDP_E_ARG, "gdp_number_edges: bad arguments");
  if(gdp::edges_on_device())
    return gdp::number_edges_device(nelem, etype, conn, nvisit, visit, edges_out, edge_nodes,
                                    max_edges);
  return gdp::number_edges(nelem, etype, conn, nvisit, visit, edges_out, edge_nodes,
                           max_edges);
}

extern "C" int gdp_rcmk(int n, const int *ia, const int *ja, int *perm, int *rperm)
{
  if(n < 0 || !ia || !ja || !perm || !rperm) return gdp::set_error(GDP_E_ARG, "gdp_rcmk: bad arguments");
  std::vector<int> p, r;
  int nlev = gdp::rcmk(n, ia, ja, p, r);
  std::copy(p.begin(), p.end(), perm);
  std::copy(r.begin(), r.end(), rperm);
  return nlev;
}
