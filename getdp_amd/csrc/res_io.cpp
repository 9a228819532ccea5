// Result files: GetDP's .res solution blocks and the vector I/O of the LinAlg
// backend they are made of.
//
//   LinAlg_PrintVector / LinAlg_ScanVector (LinAlg_SPARSKIT.cpp:174-179,
//   206-211; formatted_write_vector, Sparskit.cpp:1353-1359): one "%.16g\n"
//   per entry; fscanf("%lf") per entry.
//   LinAlg_WriteVector / LinAlg_ReadVector (:191-194, :225-229): the n doubles
//   in native byte order, followed by "\n" when writing.
//   Dof_WriteFileRES0 / Dof_WriteFileRES (DofData.cpp:631-642, :741-767): the
//   $ResFormat block (format version 1.1 and the binary flag), then per
//   solution "$Solution  /* DofData #num */", "num time time_imag step"
//   (%d %.16g %.16g %d), the vector and "$EndSolution".
//   Dof_ReadFileRES (DofData.cpp:922-1010): lines are scanned for a leading
//   '$'; $ResFormat gives the version and the binary flag; after $Solution the
//   next line is "num time step" for version <= 1.0, "num time time_imag step"
//   after that, and the vector follows.
//
// The entries go through snprintf / strtod, the same conversions as printf
// and fscanf, so a written file equals the reference's byte for byte and a
// read vector equals the reference's bit for bit.  Whole blocks are formatted
// and parsed by up to 16 threads over disjoint ranges of entries.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gdp_internal.h"
#include "host_par.h"

#define RES_TRY(call)     \
  do {                    \
    int rc_ = (call);     \
    if(rc_) return rc_;   \
  } while(0)

namespace {

using gdp::parallel_ranges;
using gdp::set_error;

int vector_host(const gdp_vector *v, std::vector<double> &h)
{
  int n = 0;
  int r = gdp_vector_size(v, &n);
  if(r) return r;
  h.resize((size_t)n);
  return n ? gdp_vector_get(v, h.data(), n) : GDP_OK;
}

// "%.16g\n" per entry, one range per thread, joined in order
void format_ascii(const std::vector<double> &h, std::string &out)
{
  const size_t n = h.size();
  const int nt = gdp::worker_count(n);
  std::vector<std::string> part(nt);
  auto fmt = [&](int t) {
    const size_t a = n * t / nt, b = n * (t + 1) / nt;
    std::string &s = part[t];
    s.reserve((b - a) * 24);
    char buf[40];
    for(size_t i = a; i < b; i++) {
      const int k = snprintf(buf, sizeof(buf), "%.16g\n", h[i]);
      s.append(buf, (size_t)k);
    }
  };
  std::vector<std::thread> th;
  for(int t = 1; t < nt; t++) th.emplace_back(fmt, t);
  fmt(0);
  for(auto &x : th) x.join();
  size_t tot = 0;
  for(auto &s : part) tot += s.size();
  out.clear();
  out.reserve(tot);
  for(auto &s : part) out += s;
}

int put(FILE *f, const void *p, size_t bytes)
{
  if(bytes && fwrite(p, 1, bytes, f) != bytes) return set_error(GDP_E_ARG, "write error (%s)", strerror(errno));
  return GDP_OK;
}

int print_vector(const std::vector<double> &h, FILE *f)
{
  std::string s;
  format_ascii(h, s);
  return put(f, s.data(), s.size());
}

int write_vector(const std::vector<double> &h, FILE *f)
{
  int r = put(f, h.data(), h.size() * sizeof(double));
  return r ? r : put(f, "\n", 1);
}

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\v' || c == '\f'; }

// n whitespace-separated numbers from [p, end): token starts found in one
// pass, conversions in parallel; *after = position after the last token
int parse_ascii(const char *p, const char *end, size_t n, double *v, const char **after)
{
  std::vector<const char *> st(n);
  for(size_t i = 0; i < n; i++) {
    while(p < end && is_space(*p)) p++;
    if(p >= end) return set_error(GDP_E_ARG, "premature end of data (%zu of %zu values)", i, n);
    st[i] = p;
    while(p < end && !is_space(*p)) p++;
  }
  *after = p;
  std::vector<char> bad(n, 0);
  parallel_ranges(n, [&](size_t a, size_t b) {
    for(size_t i = a; i < b; i++) {
      char *e;
      v[i] = strtod(st[i], &e);
      bad[i] = e == st[i];
    }
  });
  for(size_t i = 0; i < n; i++)
    if(bad[i]) return set_error(GDP_E_ARG, "value %zu is not a number", i + 1);
  return GDP_OK;
}

// next line [p, eol) and the position after it
bool next_line(const char *&p, const char *end, std::string &line)
{
  if(p >= end) return false;
  const char *q = (const char *)memchr(p, '\n', (size_t)(end - p));
  const char *e = q ? q : end;
  line.assign(p, e);
  p = q ? q + 1 : end;
  return true;
}

}  // namespace

extern "C" int gdp_vector_print(const gdp_vector *v, FILE *f)
{
  if(!v || !f) return set_error(GDP_E_ARG, "gdp_vector_print: null argument");
  std::vector<double> h;
  RES_TRY(vector_host(v, h));
  return print_vector(h, f);
}

extern "C" int gdp_vector_write(const gdp_vector *v, FILE *f)
{
  if(!v || !f) return set_error(GDP_E_ARG, "gdp_vector_write: null argument");
  std::vector<double> h;
  RES_TRY(vector_host(v, h));
  return write_vector(h, f);
}

extern "C" int gdp_vector_scan(gdp_vector *v, FILE *f)
{
  if(!v || !f) return set_error(GDP_E_ARG, "gdp_vector_scan: null argument");
  int n = 0;
  RES_TRY(gdp_vector_size(v, &n));
  std::vector<double> h((size_t)n);
  for(int i = 0; i < n; i++)
    if(fscanf(f, "%lf", &h[i]) != 1) return set_error(GDP_E_ARG, "gdp_vector_scan: value %d of %d missing", i + 1, n);
  return n ? gdp_vector_set(v, h.data(), n) : GDP_OK;
}

extern "C" int gdp_vector_read(gdp_vector *v, FILE *f)
{
  if(!v || !f) return set_error(GDP_E_ARG, "gdp_vector_read: null argument");
  int n = 0;
  RES_TRY(gdp_vector_size(v, &n));
  std::vector<double> h((size_t)n);
  if(fread(h.data(), sizeof(double), (size_t)n, f) != (size_t)n)
    return set_error(GDP_E_ARG, "gdp_vector_read: fewer than %d values", n);
  return n ? gdp_vector_set(v, h.data(), n) : GDP_OK;
}

extern "C" int gdp_res_write(const char *path, int create, int binary, int num, double time, double time_imag,
                             int step, const gdp_vector *x)
{
  if(!path || !x) return set_error(GDP_E_ARG, "gdp_res_write: null argument");
  std::vector<double> h;
  RES_TRY(vector_host(x, h));
  if(create) {
    FILE *f = fopen(path, binary ? "wb" : "w");
    if(!f) return set_error(GDP_E_ARG, "cannot open '%s' (%s)", path, strerror(errno));
    fprintf(f, "$ResFormat /* getdp_amd, %s */\n", binary ? "binary" : "ascii");
    fprintf(f, "1.1 %d\n", binary ? 1 : 0);
    fprintf(f, "$EndResFormat\n");
    if(fclose(f)) return set_error(GDP_E_ARG, "write error on '%s'", path);
  }
  FILE *f = fopen(path, binary ? "ab" : "a");
  if(!f) return set_error(GDP_E_ARG, "cannot open '%s' (%s)", path, strerror(errno));
  fprintf(f, "$Solution  /* DofData #%d */\n", num);
  fprintf(f, "%d %.16g %.16g %d\n", num, time, time_imag, step);
  int r = binary ? write_vector(h, f) : print_vector(h, f);
  fprintf(f, "$EndSolution\n");
  if(fclose(f) && !r) r = set_error(GDP_E_ARG, "write error on '%s'", path);
  return r;
}

extern "C" int gdp_res_read(const char *path, int num, int which, gdp_vector *x, double *time, double *time_imag,
                            int *step, int *count)
{
  if(!path || !x) return set_error(GDP_E_ARG, "gdp_res_read: null argument");
  int n = 0;
  RES_TRY(gdp_vector_size(x, &n));
  FILE *f = fopen(path, "rb");
  if(!f) return set_error(GDP_E_ARG, "cannot open '%s' (%s)", path, strerror(errno));
  std::string buf;
  {
    std::vector<char> chunk(1 << 20);
    size_t k;
    while((k = fread(chunk.data(), 1, chunk.size(), f)) > 0) buf.append(chunk.data(), k);
    fclose(f);
  }
  const char *p = buf.data(), *end = buf.data() + buf.size();
  double version = 0.;
  int format = 0, found = 0;
  std::vector<double> h((size_t)n), pick;
  double t_pick = 0., ti_pick = 0.;
  int s_pick = 0;
  std::string line;
  for(;;) {
    bool got = false;
    while(next_line(p, end, line))
      if(!line.empty() && line[0] == '$') {
        got = true;
        break;
      }
    if(!got) break;
    if(!line.compare(1, 9, "ResFormat")) {
      if(!next_line(p, end, line) || sscanf(line.c_str(), "%lf %d", &version, &format) != 2)
        return set_error(GDP_E_ARG, "Could not read ResFormat");
    }
    else if(!line.compare(1, 8, "Solution")) {
      if(!next_line(p, end, line)) return set_error(GDP_E_ARG, "Could not read Solution");
      int dd = 0, ts = 0;
      double t = 0., ti = 0.;
      if(version <= 1.0) sscanf(line.c_str(), "%d %lf %d", &dd, &t, &ts);
      else sscanf(line.c_str(), "%d %lf %lf %d", &dd, &t, &ti, &ts);
      const bool mine = num < 0 || dd == num;
      if(format) {
        if((size_t)(end - p) < (size_t)n * sizeof(double))
          return set_error(GDP_E_ARG, "premature end of file (time step %d)", ts);
        if(mine) std::memcpy(h.data(), p, (size_t)n * sizeof(double));
        p += (size_t)n * sizeof(double);
      }
      else if(mine) {
        const char *after;
        RES_TRY(parse_ascii(p, end, (size_t)n, h.data(), &after));
        p = after;
      }
      if(mine) {
        if(which < 0 || found == which) {
          pick = h;
          t_pick = t;
          ti_pick = ti;
          s_pick = ts;
        }
        found++;
      }
    }
    // the rest of the block, up to its closing '$' line
    bool closed = false;
    while(next_line(p, end, line))
      if(!line.empty() && line[0] == '$') {
        closed = true;
        break;
      }
    if(!closed) break;
  }
  if(count) *count = found;
  if(pick.size() != (size_t)n || (which >= 0 && which >= found))
    return set_error(GDP_E_ARG, "'%s': no solution %d of DofData %d (%d found)", path, which, num, found);
  if(time) *time = t_pick;
  if(time_imag) *time_imag = ti_pick;
  if(step) *step = s_pick;
  return n ? gdp_vector_set(x, pick.data(), n) : GDP_OK;
}
