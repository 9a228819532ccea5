// Gmsh MSH 2.x mesh files, ASCII and binary (Geo_ReadFile,
// GeoData.cpp:573-1028), read into the arrays of gdp_mesh_desc.
//
// What the reference keeps: nodes with their numbers and coordinates (read
// with a correctly rounded conversion, as fscanf("%lf") does); elements with
// number, type, physical region (first tag; 1 without tags in ASCII files, 0
// in binary ones, as the reference's two branches do) and nodes, sorted by
// element number (List_Sort with fcmp_Elm).  Here node numbers are mapped to
// 1..Nn in ascending number order (the identity for the usual contiguous
// numbering), which keeps every (NumType, Entity) key order.  Point elements
// (type 15) carry no unknowns of the supported spaces and are counted, not
// kept; other types than points, lines, triangles and tetrahedra are rejected.
//
// The ASCII sections are parsed in parallel: the line starts are found first,
// then threads convert disjoint ranges of lines.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "gdp_internal.h"
#include "host_par.h"

struct gdp_mesh_data {
  std::vector<double> xyz;     // 3 per node, ascending node number
  std::vector<int> node_num;   // original node numbers, ascending
  std::vector<int> elem_num, etype, region, elementary, conn;  // conn: 4 per element
  int npoints = 0;             // point elements read and not kept
  int binary = 0;
  double version = 0.;
};

namespace {

using gdp::parallel_ranges;
using gdp::set_error;

int nodes_of_type(int t)
{
  switch(t) {
  case 15: return 1;
  case 1: return 2;
  case 2: return 3;
  case 4: return 4;
  default: return -1;
  }
}

struct Cursor {
  const char *p, *end;
  bool line(std::string &out)
  {
    if(p >= end) return false;
    const char *q = (const char *)memchr(p, '\n', end - p);
    if(!q) q = end;
    out.assign(p, q);
    p = q < end ? q + 1 : end;
    return true;
  }
};

// starts of the next n lines from p; returns the position after them
const char *line_starts(const char *p, const char *end, size_t n, std::vector<const char *> &st)
{
  st.resize(n);
  for(size_t i = 0; i < n; i++) {
    while(p < end && (*p == '\n' || *p == '\r')) p++;
    st[i] = p;
    const char *q = (const char *)memchr(p, '\n', end - p);
    p = q ? q + 1 : end;
  }
  return p;
}

bool next_long(const char *&p, long &v)
{
  char *e;
  errno = 0;
  v = strtol(p, &e, 10);
  if(e == p || errno) return false;
  p = e;
  return true;
}

bool next_double(const char *&p, double &v)
{
  char *e;
  v = strtod(p, &e);
  if(e == p) return false;
  p = e;
  return true;
}

struct RawElem {
  int num, type, region, elementary;
  int nodes[4];
};

int finish(gdp_mesh_data &M, std::vector<int> &nnum, std::vector<double> &nxyz, std::vector<RawElem> &el)
{
  // nodes in ascending number order; number -> index map
  const size_t nn = nnum.size();
  std::vector<int> ord(nn);
  std::iota(ord.begin(), ord.end(), 0);
  bool sorted = std::is_sorted(nnum.begin(), nnum.end());
  if(!sorted) std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return nnum[a] < nnum[b]; });
  M.node_num.resize(nn);
  M.xyz.resize(nn * 3);
  for(size_t i = 0; i < nn; i++) {
    M.node_num[i] = nnum[ord[i]];
    for(int c = 0; c < 3; c++) M.xyz[3 * i + c] = nxyz[3 * (size_t)ord[i] + c];
    if(i && M.node_num[i] == M.node_num[i - 1]) return set_error(GDP_E_ARG, "node %d given twice", M.node_num[i]);
  }
  const bool contiguous = nn == 0 || (M.node_num[0] == 1 && M.node_num[nn - 1] == (int)nn);
  auto index_of = [&](int num) -> int {
    if(contiguous) return (num >= 1 && num <= (int)nn) ? num : -1;
    auto it = std::lower_bound(M.node_num.begin(), M.node_num.end(), num);
    return (it != M.node_num.end() && *it == num) ? (int)(it - M.node_num.begin()) + 1 : -1;
  };
  // elements sorted by number (fcmp_Elm), points counted and dropped
  std::stable_sort(el.begin(), el.end(), [](const RawElem &a, const RawElem &b) { return a.num < b.num; });
  const size_t ne = el.size();
  M.elem_num.reserve(ne);
  M.etype.reserve(ne);
  M.region.reserve(ne);
  M.elementary.reserve(ne);
  M.conn.reserve(ne * 4);
  for(const RawElem &r : el) {
    if(r.type == 15) {
      M.npoints++;
      continue;
    }
    M.elem_num.push_back(r.num);
    M.etype.push_back(r.type);
    M.region.push_back(r.region);
    M.elementary.push_back(r.elementary);
    const int k = nodes_of_type(r.type);
    for(int j = 0; j < 4; j++) {
      int v = 0;
      if(j < k) {
        v = index_of(r.nodes[j]);
        if(v < 0) return set_error(GDP_E_ARG, "element %d: unknown node %d", r.num, r.nodes[j]);
      }
      M.conn.push_back(v);
    }
  }
  return GDP_OK;
}

int read_msh(const char *path, gdp_mesh_data &M)
{
  FILE *f = fopen(path, "rb");
  if(!f) return set_error(GDP_E_ARG, "cannot open %s", path);
  std::vector<char> buf;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf.resize((size_t)std::max(0L, sz) + 1);
  size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
  fclose(f);
  if((long)got != sz) return set_error(GDP_E_ARG, "short read of %s", path);
  buf[(size_t)sz] = 0;
  Cursor c{buf.data(), buf.data() + sz};
  std::string line;
  std::vector<int> nnum;
  std::vector<double> nxyz;
  std::vector<RawElem> el;
  bool swap = false;
  auto rd_int = [&](const char *&p, int &v) {
    if(p + 4 > c.end) return false;
    memcpy(&v, p, 4);
    if(swap) v = (int)__builtin_bswap32((unsigned)v);
    p += 4;
    return true;
  };
  auto rd_dbl = [&](const char *&p, double &v) {
    if(p + 8 > c.end) return false;
    unsigned long long u;
    memcpy(&u, p, 8);
    if(swap) u = __builtin_bswap64(u);
    memcpy(&v, &u, 8);
    p += 8;
    return true;
  };
  while(c.line(line)) {
    if(line.empty() || line[0] != '$') continue;
    if(!line.compare(0, 11, "$MeshFormat")) {
      if(!c.line(line)) return set_error(GDP_E_ARG, "%s: truncated $MeshFormat", path);
      int fmt = 0, dsz = 0;
      if(sscanf(line.c_str(), "%lf %d %d", &M.version, &fmt, &dsz) != 3)
        return set_error(GDP_E_ARG, "%s: bad $MeshFormat", path);
      if(M.version < 2.0 || M.version >= 3.0)
        return set_error(GDP_E_UNSUPPORTED, "%s: MSH version %g (2.x supported)", path, M.version);
      M.binary = fmt != 0;
      if(M.binary) {
        if(dsz != 8) return set_error(GDP_E_UNSUPPORTED, "%s: data size %d", path, dsz);
        int one;
        const char *p = c.p;
        memcpy(&one, p, 4);
        swap = one != 1;
        c.p += 4;
        while(c.p < c.end && *c.p != '\n') c.p++;
        if(c.p < c.end) c.p++;
      }
    }
    else if(!line.compare(0, 6, "$Nodes") || !line.compare(0, 4, "$NOD")) {
      if(!c.line(line)) return set_error(GDP_E_ARG, "%s: truncated $Nodes", path);
      long n = atol(line.c_str());
      if(n < 0) return set_error(GDP_E_ARG, "%s: bad node count", path);
      nnum.resize((size_t)n);
      nxyz.resize((size_t)n * 3);
      if(M.binary) {
        const char *p = c.p;
        for(long i = 0; i < n; i++) {
          int num;
          double x, y, z;
          if(!rd_int(p, num) || !rd_dbl(p, x) || !rd_dbl(p, y) || !rd_dbl(p, z))
            return set_error(GDP_E_ARG, "%s: truncated binary nodes", path);
          nnum[i] = num;
          nxyz[3 * i] = x;
          nxyz[3 * i + 1] = y;
          nxyz[3 * i + 2] = z;
        }
        c.p = p;
      }
      else {
        std::vector<const char *> st;
        c.p = line_starts(c.p, c.end, (size_t)n, st);
        std::atomic<int> bad{0};
        parallel_ranges((size_t)n, [&](size_t a, size_t b) {
          for(size_t i = a; i < b; i++) {
            const char *p = st[i];
            long num;
            double x, y, z;
            if(!next_long(p, num) || !next_double(p, x) || !next_double(p, y) || !next_double(p, z)) {
              bad = 1;
              return;
            }
            nnum[i] = (int)num;
            nxyz[3 * i] = x;
            nxyz[3 * i + 1] = y;
            nxyz[3 * i + 2] = z;
          }
        });
        if(bad) return set_error(GDP_E_ARG, "%s: bad node line", path);
      }
    }
    else if(!line.compare(0, 9, "$Elements") || !line.compare(0, 4, "$ELM")) {
      if(!c.line(line)) return set_error(GDP_E_ARG, "%s: truncated $Elements", path);
      long n = atol(line.c_str());
      if(n < 0) return set_error(GDP_E_ARG, "%s: bad element count", path);
      el.resize((size_t)n);
      if(M.binary) {
        const char *p = c.p;
        long done = 0;
        while(done < n) {
          int hdr[3];
          if(!rd_int(p, hdr[0]) || !rd_int(p, hdr[1]) || !rd_int(p, hdr[2]))
            return set_error(GDP_E_ARG, "%s: truncated element block", path);
          const int k = nodes_of_type(hdr[0]);
          if(k < 0) return set_error(GDP_E_UNSUPPORTED, "%s: element type %d", path, hdr[0]);
          if(hdr[1] < 0 || hdr[2] < 0 || done + hdr[1] > n) return set_error(GDP_E_ARG, "%s: bad element block", path);
          for(int i = 0; i < hdr[1]; i++) {
            RawElem &r = el[(size_t)done + i];
            int v;
            if(!rd_int(p, r.num)) return set_error(GDP_E_ARG, "%s: truncated elements", path);
            r.type = hdr[0];
            r.region = 0;
            r.elementary = 0;
            for(int t = 0; t < hdr[2]; t++) {
              if(!rd_int(p, v)) return set_error(GDP_E_ARG, "%s: truncated elements", path);
              if(t == 0) r.region = v;
              else if(t == 1) r.elementary = v;
            }
            for(int j = 0; j < 4; j++) r.nodes[j] = 0;
            for(int j = 0; j < k; j++)
              if(!rd_int(p, r.nodes[j])) return set_error(GDP_E_ARG, "%s: truncated elements", path);
          }
          done += hdr[1];
        }
        c.p = p;
      }
      else {
        std::vector<const char *> st;
        c.p = line_starts(c.p, c.end, (size_t)n, st);
        std::atomic<long> bad{-1};
        auto fail = [&](long i) {
          long cur = bad.load();
          while((cur < 0 || i < cur) && !bad.compare_exchange_weak(cur, i)) {
          }
        };
        parallel_ranges((size_t)n, [&](size_t a, size_t b) {
          for(size_t i = a; i < b; i++) {
            const char *p = st[i];
            long num, type, ntag, v;
            RawElem &r = el[i];
            if(!next_long(p, num) || !next_long(p, type) || !next_long(p, ntag) || ntag < 0) {
              fail((long)i);
              return;
            }
            r.num = (int)num;
            r.type = (int)type;
            r.region = 1;  // Geo_ReadFile's default without tags (ASCII)
            r.elementary = 1;
            for(long t = 0; t < ntag; t++) {
              if(!next_long(p, v)) {
                fail((long)i);
                return;
              }
              if(t == 0) r.region = (int)v;
              else if(t == 1) r.elementary = (int)v;
            }
            const int k = nodes_of_type(r.type);
            if(k < 0) {
              fail((long)i);
              return;
            }
            for(int j = 0; j < 4; j++) r.nodes[j] = 0;
            for(int j = 0; j < k; j++) {
              if(!next_long(p, v)) {
                fail((long)i);
                return;
              }
              r.nodes[j] = (int)v;
            }
          }
        });
        if(bad >= 0) {
          const RawElem &r = el[(size_t)bad.load()];
          if(nodes_of_type(r.type) < 0) return set_error(GDP_E_UNSUPPORTED, "%s: element type %d", path, r.type);
          return set_error(GDP_E_ARG, "%s: bad element line %ld", path, bad.load() + 1);
        }
      }
    }
  }
  if(M.version == 0.) return set_error(GDP_E_ARG, "%s: no $MeshFormat section", path);
  return finish(M, nnum, nxyz, el);
}

}  // namespace

extern "C" int gdp_mesh_read(const char *path, gdp_mesh_data **out)
{
  if(!path || !out) return set_error(GDP_E_ARG, "gdp_mesh_read: null argument");
  gdp_mesh_data *M = new gdp_mesh_data;
  int rc = read_msh(path, *M);
  if(rc) {
    delete M;
    return rc;
  }
  *out = M;
  return GDP_OK;
}

extern "C" int gdp_mesh_info(const gdp_mesh_data *m, int *nnodes, int *nelem, int *npoints, int *binary)
{
  if(!m) return set_error(GDP_E_ARG, "gdp_mesh_info: null mesh");
  if(nnodes) *nnodes = (int)m->node_num.size();
  if(nelem) *nelem = (int)m->etype.size();
  if(npoints) *npoints = m->npoints;
  if(binary) *binary = m->binary;
  return GDP_OK;
}

extern "C" int gdp_mesh_get(const gdp_mesh_data *m, double *xyz, int *node_num, int *elem_num, int *etype,
                            int *region, int *elementary, int *conn)
{
  if(!m) return set_error(GDP_E_ARG, "gdp_mesh_get: null mesh");
  auto cp = [](auto *dst, const auto &v) {
    if(dst && !v.empty()) memcpy(dst, v.data(), sizeof(v[0]) * v.size());
  };
  cp(xyz, m->xyz);
  cp(node_num, m->node_num);
  cp(elem_num, m->elem_num);
  cp(etype, m->etype);
  cp(region, m->region);
  cp(elementary, m->elementary);
  cp(conn, m->conn);
  return GDP_OK;
}

extern "C" int gdp_mesh_free(gdp_mesh_data *m)
{
  delete m;
  return GDP_OK;
}

extern "C" int gdp_mesh_write(const char *path, int binary, int nnodes, const double *xyz, int nelem,
                              const int *etype, const int *region, const int *conn)
{
  if(!path || nnodes < 0 || nelem < 0 || (nnodes && !xyz) || (nelem && (!etype || !region || !conn)))
    return set_error(GDP_E_ARG, "gdp_mesh_write: bad arguments");
  for(int e = 0; e < nelem; e++)
    if(nodes_of_type(etype[e]) < 0) return set_error(GDP_E_UNSUPPORTED, "gdp_mesh_write: element type %d", etype[e]);
  FILE *f = fopen(path, "wb");
  if(!f) return set_error(GDP_E_ARG, "cannot create %s", path);
  fprintf(f, "$MeshFormat\n2.2 %d 8\n", binary ? 1 : 0);
  if(binary) {
    const int one = 1;
    fwrite(&one, 4, 1, f);
    fputc('\n', f);
  }
  fprintf(f, "$EndMeshFormat\n$Nodes\n%d\n", nnodes);
  for(int i = 0; i < nnodes; i++) {
    if(binary) {
      const int num = i + 1;
      fwrite(&num, 4, 1, f);
      fwrite(xyz + 3 * (size_t)i, 8, 3, f);
    }
    else  // %.17g round-trips every double exactly
      fprintf(f, "%d %.17g %.17g %.17g\n", i + 1, xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]);
  }
  if(binary) fputc('\n', f);
  fprintf(f, "$EndNodes\n$Elements\n%d\n", nelem);
  for(int e = 0; e < nelem; e++) {
    const int k = nodes_of_type(etype[e]);
    const int *cn = conn + (size_t)e * 4;
    if(binary) {
      if(e == 0 || etype[e] != etype[e - 1]) {  // one block per run of one type
        int run = 1;
        while(e + run < nelem && etype[e + run] == etype[e]) run++;
        const int hdr[3] = {etype[e], run, 2};
        fwrite(hdr, 4, 3, f);
      }
      const int d[3] = {e + 1, region[e], region[e]};
      fwrite(d, 4, 3, f);
      fwrite(cn, 4, (size_t)k, f);
    }
    else {
      fprintf(f, "%d %d 2 %d %d", e + 1, etype[e], region[e], region[e]);
      for(int j = 0; j < k; j++) fprintf(f, " %d", cn[j]);
      fputc('\n', f);
    }
  }
  if(binary) fputc('\n', f);
  fprintf(f, "$EndElements\n");
  const bool ok = !ferror(f);
  if(fclose(f) != 0 || !ok) return set_error(GDP_E_ARG, "write error on %s", path);
  return GDP_OK;
}
