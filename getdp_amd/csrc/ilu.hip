// ILU(0) on the device: factorisation and the two triangular sweeps.
//
// Storage: the factors live in the CSR pattern of the (renumbered) matrix:
// strictly-lower entries hold L (unit diagonal implied), the diagonal entry
// holds the inverted pivot, strictly-upper entries hold U -- the same
// information as the reference's MSR arrays alu/jlu/ju (ilut.F:1818-1941),
// in the matrix's own layout so that no second index structure is needed.
// Factor values are stored in T = float (the reference's default
// HAVE_ILU_FLOAT, CMakeLists.txt:552) or double.
//
// Factorisation: rows are grouped into dependency levels (host analysis);
// one launch per level processes its rows in parallel (a wave per row), each
// row receiving exactly the reference's sequence of operations (compiled
// without contraction), so the factors are bit-identical to the reference's.
//
// Sweeps (lusol, ilut.F:2297-2347): one lane (or a group of 2-8 lanes,
// chosen from the mean number of entries per row) per row, rows taken in
// dependency order; a row reads the solution entries it depends on as soon as
// they are published.  Outputs are pre-filled with a signalling-NaN sentinel
// that arithmetic never produces; a published entry is any non-sentinel
// value, read and written with agent-scope atomics (one 64-bit access carries
// both the value and its readiness).  Every wait is bounded by a wall-clock
// limit: on expiry the kernel records an error flag and exits.
#include <cstdlib>

#include "ilu.h"

namespace gdp {

namespace {

// ---- factorisation -----------------------------------------------------------

template <typename T> __device__ inline double prod_ll(T a, T b);
template <> __device__ inline double prod_ll<float>(float a, float b)
{
  float p = a * b;  // real*4 * real*4 is formed in single precision
  return (double)p;
}
template <> __device__ inline double prod_ll<double>(double a, double b) { return a * b; }

template <typename T>
__global__ void __launch_bounds__(256) k_ilu_copy(long nnz, const double *__restrict__ a,
                                                  T *__restrict__ lu)
{
  long k = (long)blockIdx.x * 256 + threadIdx.x;
  if(k < nnz) lu[k] = (T)a[k];
}

template <typename T>
__global__ void __launch_bounds__(128) k_ilu_level(int nrows, const int *__restrict__ rows,
                                                   const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const int *__restrict__ diag,
                                                   T *__restrict__ lu, int *__restrict__ pivot_err)
{
  int t = blockIdx.x * 128 + threadIdx.x;
  if(t >= nrows) return;
  const int i = rows[t];
  const int rs = ia[i], re = ia[i + 1], di = diag[i];
  for(int k = rs; k < di; k++) {
    const int jrow = ja[k];
    const double tl = prod_ll<T>(lu[k], lu[diag[jrow]]);
    lu[k] = (T)tl;
    // subtract tl * U(jrow, :) from the matching entries of row i
    int q = k + 1;
    const int je = ia[jrow + 1];
    for(int jj = diag[jrow] + 1; jj < je; jj++) {
      const int c = ja[jj];
      while(q < re && ja[q] < c) q++;
      if(q >= re) break;
      if(ja[q] == c) lu[q] = (T)((double)lu[q] - tl * (double)lu[jj]);
    }
  }
  const T piv = lu[di];
  if(piv == (T)0) {
    atomicMin(pivot_err, i + 1);
    return;
  }
  lu[di] = (T)(1.0 / (double)piv);
}

// The same row operation with one wave per row: the row (columns and values)
// is staged in LDS; for each strictly-lower entry k in order, lane 0's
// multiplier tl is formed as above, then the lanes take the entries of
// U(jrow) in parallel, each locating its column in the row by binary search
// and updating that entry.  Distinct U(jrow) entries hit distinct columns, and
// a barrier separates consecutive k, so every entry receives exactly the
// reference's updates in the reference's order.  Rows up to `cap` entries.
template <typename T>
__global__ void __launch_bounds__(64) k_ilu_row(int nrows, const int *__restrict__ rows,
                                                const int *__restrict__ ia, const int *__restrict__ ja,
                                                const int *__restrict__ diag, T *__restrict__ lu,
                                                int *__restrict__ pivot_err, int cap)
{
  extern __shared__ unsigned char smem[];
  int *cols = (int *)smem;
  T *vals = (T *)(smem + ((sizeof(int) * cap + 7) & ~(size_t)7));
  __shared__ double s_tl;
  const int lane = threadIdx.x;
  const int i = rows[blockIdx.x];
  const int rs = ia[i], re = ia[i + 1], di = diag[i], len = re - rs;
  for(int p = lane; p < len; p += 64) {
    cols[p] = ja[rs + p];
    vals[p] = lu[rs + p];
  }
  __syncthreads();
  for(int k = 0; k < di - rs; k++) {
    const int jrow = cols[k];
    const int dj = diag[jrow];
    if(lane == 0) {
      const double tl = prod_ll<T>(vals[k], lu[dj]);
      vals[k] = (T)tl;
      s_tl = tl;
    }
    __syncthreads();
    const double tl = s_tl;
    const int je = ia[jrow + 1];
    for(int jj = dj + 1 + lane; jj < je; jj += 64) {
      const int c = ja[jj];
      int lo = k + 1, hi = len;  // first position with cols >= c
      while(lo < hi) {
        const int mid = (lo + hi) >> 1;
        if(cols[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      if(lo < len && cols[lo] == c) vals[lo] = (T)((double)vals[lo] - tl * (double)lu[jj]);
    }
    __syncthreads();
  }
  const int dp = di - rs;
  if(lane == 0) {
    const T piv = vals[dp];
    if(piv == (T)0) atomicMin(pivot_err, i + 1);
    else vals[dp] = (T)(1.0 / (double)piv);
  }
  __syncthreads();
  for(int p = lane; p < len; p += 64) lu[rs + p] = vals[p];
}

// ---- sweeps -----------------------------------------------------------------------
template <typename T> struct Sentinel;
template <> struct Sentinel<double> {
  using U = unsigned long long;
  static constexpr U bits = 0x7FF0DEADBEEF0001ULL;  // signalling NaN
};
template <> struct Sentinel<float> {
  using U = unsigned int;
  static constexpr U bits = 0x7F80BEEFu;  // signalling NaN
};

template <typename T, int SCOPE = __HIP_MEMORY_SCOPE_AGENT> __device__ inline typename Sentinel<T>::U ld_bits(const T *p)
{
  using U = typename Sentinel<T>::U;
  return __hip_atomic_load((const U *)p, __ATOMIC_RELAXED, SCOPE);
}
template <typename T, int SCOPE = __HIP_MEMORY_SCOPE_AGENT> __device__ inline void st_val(T *p, T v)
{
  using U = typename Sentinel<T>::U;
  U b;
  __builtin_memcpy(&b, &v, sizeof(T));
  __hip_atomic_store((U *)p, b, __ATOMIC_RELAXED, SCOPE);
}
template <typename T> __device__ inline T as_val(typename Sentinel<T>::U b)
{
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}

// Kernels enqueued inside a Krylov cycle are skipped once the solver's
// control block says the solve (gmode 1) or the cycle (gmode 2) has ended;
// gate[0] is the done flag, gate[3] the end-of-cycle flag.
__device__ inline bool is_gated(const int *gate, int gmode)
{
  if(!gmode) return false;
  if(gmode == 1) return gate[0] != 0;
  return (gate[0] | gate[3]) != 0;
}

// Bounded wait: a thread may spend at most tlimit wall-clock ticks waiting
// in one sweep; once any thread has given up, every other waiter gives up
// too, so a sweep always ends.
__device__ inline bool wait_expired(unsigned long long &t0, unsigned long long tlimit, int *err)
{
  unsigned long long now = wall_clock64();
  if(!t0) {
    t0 = now;
    return false;
  }
  if(now - t0 > tlimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    atomicOr(err, 1);
    return true;
  }
  return false;
}

// Sweeps, one kernel for both directions: rows in dependency order
// (ascending for L, descending for U) over consecutive waves and blocks, so a
// row only waits on rows of earlier waves, which the in-order dispatch of
// workgroups has already made resident.  G lanes serve one row (G = 1 .. 32,
// chosen from the mean number of entries per row): they wait for G of its
// dependencies at once (one memory round trip instead of G on the critical
// path), then every lane of the row subtracts the G products in ascending
// column position -- lusol's order and rounding, whatever G.  The products
// travel by cross-lane reads executed by the whole wave, outside divergent
// branches.  The forward sweep fills `x` (the backward sweep's output) with
// the sentinel; the backward sweep resets y to the sentinel as it consumes it.
template <typename TV> __device__ inline TV lane_read(TV v, int src) { return __shfl(v, src, kWave); }

// Rows base .. base + 64/G - 1 (in dependency order) by one wave; SCOPE is
// the memory scope of the publications and polls.
template <typename TF, typename TV, int G, int C, bool FWD, int SCOPE>
__device__ inline void sweep_rows(int base, int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                  const int *__restrict__ diag, const TF *__restrict__ lu,
                                  const TV *__restrict__ r, TV scale, TV *y, TV *x, int *err,
                                  unsigned long long tlimit, TV *addto)
{
  constexpr int R = kWave / G;  // rows per wave
  constexpr int S = G * C;      // entries of a row in flight at once
  constexpr unsigned long long GM = (G == 64) ? ~0ull : ((1ull << G) - 1);
  const int lane = threadIdx.x & (kWave - 1), q = lane & (G - 1), gbase = lane - q;
  TV *out = FWD ? y : x;
  {
    // rows t = base .. base + R - 1 of this wave, in dependency order
    const int t = base + lane / G;
    const bool valid = t < n;
    const int i = FWD ? t : n - 1 - t;
    TV acc = 0;
    int cb = 0, kend = 0, di = 0;
    if(valid) {
      di = diag[i];
      if(FWD) {
        acc = r[i] * scale;
        cb = ia[i];
        kend = di;
        if(q == 0) st_val<TV, SCOPE>(x + i, as_val<TV>(Sentinel<TV>::bits));
      }
      else {
        acc = y[i];
        cb = di + 1;
        kend = ia[i + 1];
      }
    }
    // every lane of the row has read y[i] (one instruction) before lane 0
    // resets it for the next forward sweep
    if(!FWD && valid && q == 0) st_val<TV, SCOPE>(y + i, as_val<TV>(Sentinel<TV>::bits));
    // a lane holds positions cb + q + G*c, c < C, of the row's current span of
    // S entries; their columns and factor entries are loaded when the span
    // starts, off the dependency chain of the polls
    const TV ldi = (!FWD && valid) ? (TV)lu[di] : (TV)0;
    int jk[C];
    TV lk[C], prod[C];
    unsigned pend = 0;  // bit c: position c still waits for its dependency
    auto fetch = [&]() {
      pend = 0;
#pragma unroll
      for(int c = 0; c < C; c++) {
        const int k = cb + q + G * c;
        prod[c] = 0;
        if(k < kend) {
          jk[c] = ja[k];
          lk[c] = (TV)lu[k];
          pend |= 1u << c;
        }
      }
    };
    if(valid) fetch();
    auto publish = [&]() {
      if(q == 0) {
        if(FWD) st_val<TV, SCOPE>(y + i, acc);
        else {
          const TV xi = ldi * acc;
          st_val<TV, SCOPE>(x + i, xi);
          if(addto) addto[i] = addto[i] + xi;  // fused z += x of the caller
        }
      }
    };
    bool done = !valid;
    if(valid && cb >= kend) {
      publish();
      done = true;
    }
    unsigned long long t0 = 0;
    for(;;) {
      if(!done && pend) {
        bool any_wait = false;
#pragma unroll
        for(int c = 0; c < C; c++) {
          if(pend & (1u << c)) {
            auto b = ld_bits<TV, SCOPE>(out + jk[c]);
            if(b != Sentinel<TV>::bits) {
              prod[c] = lk[c] * as_val<TV>(b);
              pend &= ~(1u << c);
            }
            else
              any_wait = true;
          }
        }
        if(any_wait) {
          if(wait_expired(t0, tlimit, err)) pend = 0;
          else __builtin_amdgcn_s_sleep(1);
        }
      }
      const unsigned long long m = __ballot(!pend || done);
      const bool ready = !done && ((m >> gbase) & GM) == GM;
      if(G == 1) {
#pragma unroll
        for(int c = 0; c < C; c++)
          if(ready && cb + c < kend) acc = acc - prod[c];
      }
      else if(__any(ready)) {  // wave-uniform: cross-lane reads only when a row can use them
#pragma unroll
        for(int c = 0; c < C; c++) {
#pragma unroll
          for(int j = 0; j < G; j++) {
            const TV pj = lane_read(prod[c], gbase + j);
            if(ready && cb + G * c + j < kend) acc = acc - pj;
          }
        }
      }
      if(ready) {
        cb += S;
        if(cb >= kend) {
          publish();
          done = true;
        }
        else
          fetch();
      }
      // keeps the publishing store inside the iteration (a lane waiting on a
      // row of the same wave sees it on a later iteration)
      __asm__ volatile("" ::: "memory");
      if(__all(done)) break;
    }
  }
}

template <typename TF, typename TV, int G, int C, bool FWD>
__global__ void __launch_bounds__(256) k_sweep(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                               const int *__restrict__ diag, const TF *__restrict__ lu,
                                               const TV *__restrict__ r, TV scale, TV *y, TV *x, int *err,
                                               unsigned long long tlimit, const int *gate, int gmode,
                                               TV *addto)
{
  if(is_gated(gate, gmode)) return;  // uniform across the grid
  const int base = (blockIdx.x * (256 / kWave) + (int)(threadIdx.x / kWave)) * (kWave / G);
  sweep_rows<TF, TV, G, C, FWD, __HIP_MEMORY_SCOPE_AGENT>(base, n, ia, ja, diag, lu, r, scale, y, x, err, tlimit,
                                                          addto);
}

template <typename T> __global__ void k_fill_bits(long n, T *p)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i < n) p[i] = as_val<T>(Sentinel<T>::bits);
}

__global__ void k_diag_pos(int n, const int *ia, const int *ja, int *diag, int *missing)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if(i >= n) return;
  int lo = ia[i], hi = ia[i + 1];
  while(lo < hi) {
    int mid = (lo + hi) >> 1;
    if(ja[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  if(lo < ia[i + 1] && ja[lo] == i) diag[i] = lo;
  else {
    diag[i] = lo;
    atomicMin(missing, i + 1);
  }
}

// ILUTP's column permutation of the factors: x[j] = t[colmap[j]] (and
// addto[j] += x[j] when given).
template <typename TV>
__global__ void __launch_bounds__(256) k_colperm(int n, const TV *__restrict__ t, const int *__restrict__ colmap,
                                                 TV *__restrict__ x, TV *addto, const int *gate, int gmode)
{
  if(is_gated(gate, gmode)) return;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if(j >= n) return;
  const TV v = t[colmap[j]];
  x[j] = v;
  if(addto) addto[j] = addto[j] + v;
}

}  // namespace

// ---- host side ------------------------------------------------------------------------

void IluFactor::release()
{
  dfree(lu32);
  dfree(lu64);
  dfree(diag);
  dfree(level_rows);
  dfree(tmp64);
  dfree(tmp32);
  dfree(err);
  dfree(own_ia);
  dfree(own_ja);
  dfree(colmap);
  dfree(pbuf64);
  dfree(pbuf32);
  ia = ja = nullptr;
  level_ptr.clear();
  n = 0;
}

// Levels, lanes per row and the sweep buffers for the host pattern hia/hja
// (strictly lower entries first in each row, strictly upper entries last).
int IluFactor::setup_sweeps(const int *hia, const int *hja)
{
  std::vector<int> lev;
  nlevels = lower_levels(n, hia, hja, lev);
  std::vector<int> cnt(nlevels + 1, 0);
  for(int i = 0; i < n; i++) cnt[lev[i] + 1]++;
  for(int l = 0; l < nlevels; l++) cnt[l + 1] += cnt[l];
  level_ptr = cnt;
  std::vector<int> rows(n);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for(int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
  std::vector<int> ulev;
  upper_nlevels = upper_levels(n, hia, hja, ulev);
  // lanes per row of each sweep: the power of two at or above the mean count
  // of strictly lower (upper) entries, 1 below 1.5 entries.  When the levels
  // hold many rows (short dependency chains: ILU(0) in RCMK order) the sweep
  // is throughput-bound and the groups stop at 8 lanes (elasticity 32^3, 20
  // entries a side: 3.4 ms per application with 8 lanes, 3.2 with 16, 3.9
  // with 32; Poisson 256^2: 2.0 ms with 8, 3.4 with 32).  When the chain is
  // long (under 64 rows a level: ILUT/ILUTP with fill) its latency dominates
  // and a group as wide as the row takes all its entries in one span
  // (ILUTP(20) on Poisson 256^2, 20 entries a side, 10363 levels: 58.8 ms
  // with 8 lanes, 52.8 with 16, 40.8 with 32).
  long nl = 0, nu = 0;
  for(int i = 0; i < n; i++)
    for(int k = hia[i]; k < hia[i + 1]; k++) {
      if(hja[k] < i) nl++;
      else if(hja[k] > i) nu++;
    }
  const int cap = (long)nlevels * 64 > n ? 32 : 8;
  auto lanes = [&](long c) {
    const double avg = n ? (double)c / n : 0.;
    if(avg < 1.5) return 1;
    int g = 2;
    while(g < avg && g < cap) g *= 2;
    return g;
  };
  maxrow = 1;
  for(int i = 0; i < n; i++) maxrow = std::max(maxrow, hia[i + 1] - hia[i]);
  g_lower = lanes(nl);
  g_upper = lanes(nu);
  if(const char *e = std::getenv("GDP_ILU_LANES")) g_lower = g_upper = std::max(1, std::min(32, atoi(e)));
  GDP_TRY(dalloc(&level_rows, (size_t)std::max(1, n)));
  GDP_TRY(h2d(level_rows, rows.data(), (size_t)n));
  GDP_TRY(dalloc(&diag, (size_t)std::max(1, n)));
  GDP_TRY(dalloc(&err, 4));
  int init[4] = {0x7fffffff, 0, 0x7fffffff, 0};
  GDP_TRY(h2d(err, init, 4));
  GDP_TRY(dalloc(&tmp64, (size_t)std::max(1, n)));
  GDP_TRY(dalloc(&tmp32, (size_t)std::max(1, n)));
  return GDP_OK;
}

int IluFactor::analyse(const DevCSR &B, const std::vector<int> &hia, const std::vector<int> &hja)
{
  release();
  n = B.n;
  nnz = B.nnz;
  ia = B.ia;
  ja = B.ja;
  GDP_TRY(setup_sweeps(hia.data(), hja.data()));
  if(n)
    hipLaunchKernelGGL(k_diag_pos, dim3((n + 255) / 256), dim3(256), 0, ctx().stream, n, ia, ja,
                       diag, err + 2);
  GDP_HIP(hipGetLastError());
  int h[4];
  GDP_TRY(d2h(h, err, 4));
  if(h[2] != 0x7fffffff)
    return set_error(GDP_E_ZERO_PIVOT, "Zero pivot on line %d in ILU (missing diagonal)", h[2]);
  GDP_TRY(reset_sweep_state());
  return GDP_OK;
}

template <typename T> int IluFactor::load(const IlutRows<T> &f, bool f64)
{
  release();
  n = (int)f.dpos.size();
  nnz = (long)f.col.size();
  GDP_TRY(setup_sweeps(f.ptr.data(), f.col.data()));
  GDP_TRY(dalloc(&own_ia, (size_t)n + 1));
  GDP_TRY(dalloc(&own_ja, (size_t)std::max(1L, nnz)));
  GDP_TRY(h2d(own_ia, f.ptr.data(), (size_t)n + 1));
  GDP_TRY(h2d(own_ja, f.col.data(), (size_t)nnz));
  GDP_TRY(h2d(diag, f.dpos.data(), (size_t)n));
  ia = own_ia;
  ja = own_ja;
  is_f64 = f64;
  if(f64) {
    std::vector<double> v(f.val.begin(), f.val.end());
    GDP_TRY(dalloc(&lu64, (size_t)std::max(1L, nnz)));
    GDP_TRY(h2d(lu64, v.data(), (size_t)nnz));
  }
  else {
    std::vector<float> v(f.val.begin(), f.val.end());
    GDP_TRY(dalloc(&lu32, (size_t)std::max(1L, nnz)));
    GDP_TRY(h2d(lu32, v.data(), (size_t)nnz));
  }
  if(!f.newpos.empty()) {
    GDP_TRY(dalloc(&colmap, (size_t)std::max(1, n)));
    GDP_TRY(h2d(colmap, f.newpos.data(), (size_t)n));
    GDP_TRY(dalloc(&pbuf64, (size_t)std::max(1, n)));
    GDP_TRY(dalloc(&pbuf32, (size_t)std::max(1, n)));
  }
  return reset_sweep_state();
}

template int IluFactor::load<float>(const IlutRows<float> &, bool);
template int IluFactor::load<double>(const IlutRows<double> &, bool);

int IluFactor::reset_sweep_state()
{
  hipStream_t s = ctx().stream;
  if(n) {
    hipLaunchKernelGGL(k_fill_bits<double>, dim3((n + 255) / 256), dim3(256), 0, s, (long)n, tmp64);
    hipLaunchKernelGGL(k_fill_bits<float>, dim3((n + 255) / 256), dim3(256), 0, s, (long)n, tmp32);
  }
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

// One launch per dependency level: a wave per row (row staged in LDS) when
// the longest row fits kRowCap entries, else a thread per row.
constexpr int kRowCap = 1024;

template <typename T> void IluFactor::enqueue_levels(T *lu)
{
  hipStream_t s = ctx().stream;
  const bool wave_rows = maxrow <= kRowCap && !std::getenv("GDP_ILU_THREAD_ROWS");
  const size_t lds = (((size_t)maxrow * sizeof(int) + 7) & ~(size_t)7) + (size_t)maxrow * sizeof(T);
  for(int l = 0; l < nlevels; l++) {
    const int cnt = level_ptr[l + 1] - level_ptr[l];
    if(!cnt) continue;
    if(wave_rows)
      hipLaunchKernelGGL(k_ilu_row<T>, dim3(cnt), dim3(64), lds, s, cnt, level_rows + level_ptr[l], ia, ja,
                         diag, lu, err, maxrow);
    else
      hipLaunchKernelGGL(k_ilu_level<T>, dim3((cnt + 127) / 128), dim3(128), 0, s, cnt,
                         level_rows + level_ptr[l], ia, ja, diag, lu, err);
  }
}

int IluFactor::factor(const double *a, bool f64, int *zero_pivot_row)
{
  hipStream_t s = ctx().stream;
  is_f64 = f64;
  dfree(lu32);
  dfree(lu64);
  if(f64) GDP_TRY(dalloc(&lu64, (size_t)std::max<long>(1, nnz)));
  else GDP_TRY(dalloc(&lu32, (size_t)std::max<long>(1, nnz)));
  int init = 0x7fffffff;
  GDP_TRY(h2d(err, &init, 1));
  dim3 g((unsigned)((nnz + 255) / 256));
  if(nnz) {
    if(f64) hipLaunchKernelGGL(k_ilu_copy<double>, g, dim3(256), 0, s, nnz, a, lu64);
    else hipLaunchKernelGGL(k_ilu_copy<float>, g, dim3(256), 0, s, nnz, a, lu32);
  }
  if(f64) enqueue_levels(lu64);
  else enqueue_levels(lu32);
  GDP_HIP(hipGetLastError());
  int h;
  GDP_TRY(d2h(&h, err, 1));
  *zero_pivot_row = (h == 0x7fffffff) ? 0 : h;
  int zero = 0;
  GDP_TRY(h2d(err + 1, &zero, 1));
  return GDP_OK;
}

namespace {
__global__ void k_set_int(int *p, int v) { *p = v; }
}  // namespace

int IluFactor::reserve(bool f64)
{
  is_f64 = f64;
  dfree(lu32);
  dfree(lu64);
  if(f64) GDP_TRY(dalloc(&lu64, (size_t)std::max<long>(1, nnz)));
  else GDP_TRY(dalloc(&lu32, (size_t)std::max<long>(1, nnz)));
  return GDP_OK;
}

int IluFactor::factor_enqueue(const double *a)
{
  hipStream_t s = ctx().stream;
  if(!(is_f64 ? (void *)lu64 : (void *)lu32)) return set_error(GDP_E_STATE, "ILU storage not reserved");
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, err, 0x7fffffff);
  dim3 g((unsigned)((nnz + 255) / 256));
  if(nnz) {
    if(is_f64) hipLaunchKernelGGL(k_ilu_copy<double>, g, dim3(256), 0, s, nnz, a, lu64);
    else hipLaunchKernelGGL(k_ilu_copy<float>, g, dim3(256), 0, s, nnz, a, lu32);
  }
  if(is_f64) enqueue_levels(lu64);
  else enqueue_levels(lu32);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int IluFactor::factor_status(int *zero_pivot_row)
{
  int h;
  GDP_TRY(d2h(&h, err, 1));
  *zero_pivot_row = (h == 0x7fffffff) ? 0 : h;
  return GDP_OK;
}

unsigned long long IluFactor::wait_limit() const
{
  return 100000000ull;  // wall_clock64 runs at 100 MHz: 1 s per dependency wait
}

namespace {

template <typename TF, typename TV, bool FWD>
void sweep(int G, int n, const int *ia, const int *ja, const int *diag, const TF *lu, const TV *r,
           TV scale, TV *y, TV *x, int *err, unsigned long long tl, const int *gate, int gmode, TV *addto)
{
  hipStream_t s = ctx().stream;
  // one entry per lane in flight (C = 1): spans of 2-4 entries per lane were
  // measured no faster (the chain's cost is the publication latency, not the
  // number of polls)
  if(G != 2 && G != 4 && G != 8 && G != 16 && G != 32) G = 1;
#define GDP_SWEEP(GG)                                                                                  \
  hipLaunchKernelGGL((k_sweep<TF, TV, GG, 1, FWD>), dim3((unsigned)(((long)n * GG + 255) / 256)), dim3(256), 0, \
                     s, n, ia, ja, diag, lu, r, scale, y, x, err, tl, gate, gmode, addto)
  switch(G) {
  case 2: GDP_SWEEP(2); break;
  case 4: GDP_SWEEP(4); break;
  case 8: GDP_SWEEP(8); break;
  case 16: GDP_SWEEP(16); break;
  case 32: GDP_SWEEP(32); break;
  default: GDP_SWEEP(1); break;
  }
#undef GDP_SWEEP
}

}  // namespace

template <typename TV>
int IluFactor::apply(const TV *r, TV scale, TV *x, const int *gate, int gmode, TV *addto)
{
  hipStream_t s = ctx().stream;
  if(!n) return GDP_OK;
  TV *t, *out = x, *add = addto;
  if constexpr(sizeof(TV) == 8) {
    t = tmp64;
    if(colmap) out = pbuf64;
  }
  else {
    t = tmp32;
    if(colmap) out = pbuf32;
  }
  if(colmap) add = nullptr;
  if(is_f64) {
    sweep<double, TV, true>(g_lower, n, ia, ja, diag, lu64, r, scale, t, out, err + 1, wait_limit(),
                            gate, gmode, nullptr);
    sweep<double, TV, false>(g_upper, n, ia, ja, diag, lu64, nullptr, scale, t, out, err + 1,
                             wait_limit(), gate, gmode, add);
  }
  else {
    sweep<float, TV, true>(g_lower, n, ia, ja, diag, lu32, r, scale, t, out, err + 1, wait_limit(),
                           gate, gmode, nullptr);
    sweep<float, TV, false>(g_upper, n, ia, ja, diag, lu32, nullptr, scale, t, out, err + 1,
                            wait_limit(), gate, gmode, add);
  }
  if(colmap)
    hipLaunchKernelGGL(k_colperm<TV>, dim3((n + 255) / 256), dim3(256), 0, s, n, out, colmap, x, addto, gate,
                       gmode);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int IluFactor::check_timeout()
{
  int h = 0;
  GDP_TRY(d2h(&h, err + 1, 1));
  if(h) return set_error(GDP_E_TIMEOUT, "ILU sweep: a dependency wait exceeded its limit");
  return GDP_OK;
}

template int IluFactor::apply<double>(const double *, double, double *, const int *, int, double *);
template int IluFactor::apply<float>(const float *, float, float *, const int *, int, float *);

}  // namespace gdp
