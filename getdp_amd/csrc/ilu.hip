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
#include <cstring>

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
// too, so a sweep always ends.  The clock and the error flag are read every
// 64th failed poll only: each is a round trip of its own (a scalar memory
// read, a device-scope load) that would otherwise lengthen every poll.
__device__ inline bool wait_expired(unsigned long long &t0, unsigned &spins, unsigned long long tlimit, int *err)
{
  if(++spins & 63) return false;
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
    // S entries; their columns and factor entries are loaded a span ahead
    // (while the current span's dependencies are polled), so that moving on
    // to the next span does not wait for a load from memory: with factors
    // larger than the Infinity Cache that was an HBM latency on the chain
    const TV ldi = (!FWD && valid) ? (TV)lu[di] : (TV)0;
    int jk[C], jn[C];
    TV lk[C], ln[C], prod[C];
    unsigned pend = 0;  // bit c: position c still waits for its dependency
    auto load_next = [&]() {
#pragma unroll
      for(int c = 0; c < C; c++) {
        const int k = cb + S + q + G * c;
        if(k < kend) {
          jn[c] = ja[k];
          ln[c] = (TV)lu[k];
        }
      }
    };
    // first span: loaded in place; later spans: taken from the prefetch
    if(valid) {
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
      load_next();
    }
    auto fetch = [&]() {
      pend = 0;
#pragma unroll
      for(int c = 0; c < C; c++) {
        const int k = cb + q + G * c;
        prod[c] = 0;
        if(k < kend) {
          jk[c] = jn[c];
          lk[c] = ln[c];
          pend |= 1u << c;
        }
      }
      load_next();
    };
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
    unsigned spins = 0;
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
          if(wait_expired(t0, spins, tlimit, err)) pend = 0;
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

// ---- level walks in one workgroup ----------------------------------------------
//
// With a long dependency chain of narrow levels (threshold ILU with fill:
// ILUTP(20) on Poisson 256^2 has about 10,000 levels of 6 rows per triangle)
// the sweeps above pay one publication round trip through the memory
// hierarchy per level.  Here one workgroup walks the levels in order instead.
// A chunk holds up to 64/G rows of one level, G lanes per row, lane q of a
// row holding its entries q, q + G, ...; a row's stream position is its index
// in level order, and the solution entries of the last kRing stream
// positions stay in an LDS ring indexed by position mod kRing (the host
// checks that every dependency is still there when it is read: in level
// order a dependency is a few levels back, although the row numbers of one
// level span the whole matrix); the chunks are loaded a few chunks ahead.  A level then costs LDS reads, the row's subtraction
// chain and an LDS write, with no memory round trip.  Every row subtracts its
// products in ascending position (lusol's order; padding subtracts +0 last).
constexpr int kRing = 16384;
// One lane per row, up to 32 entries: the row's products and its
// subtraction chain stay in one lane, with no cross-lane reads.  The entry
// slots of a sweep are the first of these that holds its longest row.
constexpr int kWaveSlots[] = {4, 8, 16, 24, 32};
constexpr int kNumWaveShapes = 5;

// Chunk records.  Entry slots are stored as 16-byte records, lane-interleaved
// (record j of lane l in chunk c at (c * NL + j) * 64 + l): two (column,
// value) pairs per record with FP32 factors, one (column, -, value) with FP64
// ones, the column as a byte offset into an FP64 ring (no shift in the walk), so a chunk is NL = slots / 2 (or slots) coalesced 16-byte loads per
// lane, plus one 16-byte load of the row's fields (row, forward output
// position, ring slot, need), the right-hand side and, backward, the
// inverted pivot.  A wave keeps NB chunks in registers and rotates through
// them without copies: the loads for its chunk k + NB - 1 are issued before
// chunk k is computed, so a chunk's loads have NB - 1 chunks of the wave's
// own work (and the team's in between) to arrive.  NB is 4 while the loads
// in flight stay within the 63 the wave's memory counter tracks, else 2.
__host__ __device__ constexpr int walk_nl(bool f64, int wg) { return f64 ? wg : wg / 2; }
__host__ __device__ constexpr int walk_nb(bool f64, int wg) { return walk_nl(f64, wg) + 4 <= 20 ? 4 : 2; }
// prefetch padding per wave (chunks) after the last step: the largest NB
constexpr int kWalkPad = 4;

// Several waves of one workgroup may share the walk (a team): wave w takes
// chunks w, w + T, w + 2T, ... (T waves).  Chunk c may start once every chunk
// below need[c] is done: need[c] covers the chunks holding c's dependencies,
// and, for each ring slot c writes, the chunk that last wrote it and the
// last chunk that reads the value it holds (the host computes need from the
// sequential order, in which the ring check holds).  Each wave publishes how
// many of its chunks are done in LDS; chunks below s are all done when every
// wave v has finished its ceil((s - v) / T) chunks below s.  At most four
// waves (one per SIMD), so that each may hold 512 registers.
constexpr int kMaxTeam = 4;

// nch chunks (a multiple of team * kWalkPad), followed by team * kWalkPad
// padding chunks that the last prefetches read; padding rows are -1 and never
// stored.
template <typename TF, typename TV, int WG, bool FWD>
__global__ void __launch_bounds__(kWave *kMaxTeam) k_wave_sweep(int nch, const uint4 *__restrict__ ent,
                                                               const int4 *__restrict__ meta,
                                                               const TF *__restrict__ dinv, const TV *__restrict__ pre,
                                                               TV *out, int *err, unsigned long long tlimit,
                                                               const int *gate, int gmode)
{
  if(is_gated(gate, gmode)) return;
  constexpr bool F64 = sizeof(TF) == 8;
  constexpr int NL = walk_nl(F64, WG), NB = walk_nb(F64, WG);
  struct Rec {
    uint4 e[NL];
    int4 m;  // row, forward output position, ring slot, need
    TV p;
    TF d;
  };
  __shared__ TV ring[kRing + 1];
  __shared__ int prog[kMaxTeam];
  const int lane = threadIdx.x & (kWave - 1);
  const int team = blockDim.x / kWave, wv = threadIdx.x / kWave;
  if(threadIdx.x == 0) ring[kRing] = 0;  // the padding slot
  if(threadIdx.x < kMaxTeam) prog[threadIdx.x] = 0;
  __syncthreads();
  bool expired = false;
  auto load = [&](Rec &b, int c) {
#pragma unroll
    for(int j = 0; j < NL; j++) b.e[j] = ent[((size_t)c * NL + j) * kWave + lane];
    const size_t li = (size_t)c * kWave + lane;
    b.m = meta[li];
    b.p = pre[li];
    if constexpr(!FWD) b.d = dinv[li];
  };
  auto col = [&](const Rec &b, int w) -> int {
    if constexpr(F64) return (int)b.e[w].x;
    else return (int)((w & 1) ? b.e[w >> 1].z : b.e[w >> 1].x);
  };
  auto val = [&](const Rec &b, int w) -> TF {
    if constexpr(F64) return __hiloint2double((int)b.e[w].w, (int)b.e[w].z);
    else return __uint_as_float((w & 1) ? b.e[w >> 1].w : b.e[w >> 1].y);
  };
  // this wave's k-th chunk
  auto chunk = [&](int k) { return wv + team * k; };
  auto process = [&](const Rec &rc, int k) {
    if(team > 1 && rc.m.w > 0) {
      // wait until every chunk below rc.m.w is done (bounded); the clock is
      // read once 256 polls have failed, and every 256th poll after that (a
      // read of it is a scalar memory round trip)
      const int s = rc.m.w;
      const int want = lane < team && s > lane ? (s - lane + team - 1) / team : 0;
      unsigned long long t0 = 0;
      unsigned spins = 0;
      while(true) {
        const int have =
          lane < team ? __hip_atomic_load(prog + lane, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
        if(__all(have >= want) || expired) break;
        if((++spins & 255) == 0) {
          const unsigned long long now = wall_clock64();
          if(!t0) t0 = now;
          else if(now - t0 > tlimit) {
            expired = true;
            if(lane == 0) atomicOr(err, 1);
          }
        }
      }
    }
    // every ring read is issued before the first product: left to itself
    // the scheduler interleaves them in groups of four, each waiting out a
    // full LDS round trip on the row's chain
    TV rv[WG], prod[WG];
#pragma unroll
    for(int w = 0; w < WG; w++)
      rv[w] = sizeof(TV) == 8 ? *(const TV *)((const unsigned char *)ring + col(rc, w)) : ring[col(rc, w) >> 3];
    __builtin_amdgcn_sched_group_barrier(0x100, WG, 0);      // the DS reads first
    __builtin_amdgcn_sched_group_barrier(0x002, 4 * WG, 0);  // then the arithmetic
#pragma unroll
    for(int w = 0; w < WG; w++) prod[w] = (TV)val(rc, w) * rv[w];
    // and every product before the subtraction chain, which then waits on
    // one subtraction per entry instead of a product and a subtraction
#pragma unroll
    for(int w = 0; w < WG; w++) __asm__ volatile("" : "+v"(prod[w]));
    TV acc = rc.p;
#pragma unroll
    for(int w = 0; w < WG; w++) acc = acc - prod[w];  // lusol's order; padding subtracts +0 last
    if(rc.m.x >= 0) {
      const TV xv = FWD ? acc : (TV)rc.d * acc;
      ring[rc.m.z] = xv;  // the slot of its stream position
      if constexpr(FWD) out[rc.m.y] = xv;  // the backward sweep's input, in its order
      else out[rc.m.x] = xv;
    }
    if(team > 1 && lane == 0)  // this wave's chunks up to k are done
      __hip_atomic_store(prog + wv, k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const int nk = nch / team;
  Rec b[NB];
#pragma unroll
  for(int d = 0; d + 1 < NB; d++) load(b[d], chunk(d));
  for(int k0 = 0; k0 < nk; k0 += NB) {
#pragma unroll
    for(int d = 0; d < NB; d++) {
      load(b[(d + NB - 1) % NB], chunk(k0 + d + NB - 1));
      process(b[d], k0 + d);
    }
  }
}

// addto += x after the one-wave backward sweep (the same single addition the
// sync-free sweep fuses; inside the sweep its load would sit on the chain)
template <typename TV>
__global__ void __launch_bounds__(256) k_wave_addto(int n, const TV *__restrict__ x, TV *__restrict__ addto,
                                                    const int *gate, int gmode)
{
  if(is_gated(gate, gmode)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n) addto[i] = addto[i] + x[i];
}

// pre[p] = scale * r[row[p]] in the forward stream's order
template <typename TV>
__global__ void __launch_bounds__(256) k_wave_gather(long cnt, const int *__restrict__ row, const TV *__restrict__ r,
                                                     TV scale, TV *__restrict__ pre, const int *gate, int gmode)
{
  if(is_gated(gate, gmode)) return;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if(p < cnt) {
    const int i = row[p];
    pre[p] = i >= 0 ? r[i] * scale : (TV)0;
  }
}

// val[p] = lu[kpos[p]] (0 for padding), after each factorisation
template <typename TF>
__global__ void __launch_bounds__(256) k_wave_pack(long cnt, const int *__restrict__ kpos, const TF *__restrict__ lu,
                                                   TF *__restrict__ val)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if(p < cnt) {
    const int k = kpos[p];
    val[p] = k >= 0 ? lu[k] : (TF)0;
  }
}

// The walk's 16-byte entry records from the slot arrays (slot w of lane l in
// chunk c at ((c * wg + w) * 64 + l)): column and value bits, after each
// factorisation (enqueued: usable under capture).
template <typename TF>
__global__ void __launch_bounds__(256) k_wave_pack_ent(long cnt, int wg, const int *__restrict__ rk0,
                                                       const int *__restrict__ rcnt, const int *__restrict__ ecol,
                                                       const TF *__restrict__ lu, unsigned *__restrict__ ent)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if(p >= cnt) return;
  const long c = p / ((long)wg * kWave);
  const int w = (int)(p / kWave % wg), l = (int)(p % kWave);
  const size_t q = (size_t)c * kWave + l;
  const bool used = w < rcnt[q];  // padding: the zero ring slot and a zero value
  const int k = used ? rk0[q] + w : -1;
  const TF v = used ? lu[k] : (TF)0;
  const unsigned off = (unsigned)(used ? ecol[k] : kRing) * 8u;  // byte offset into an FP64 ring
  if constexpr(sizeof(TF) == 4) {
    const size_t word = (((size_t)c * (wg / 2) + w / 2) * kWave + l) * 4 + (w & 1) * 2;
    ent[word] = off;
    ent[word + 1] = __float_as_uint(v);
  }
  else {
    const size_t word = (((size_t)c * wg + w) * kWave + l) * 4;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    ent[word] = off;
    ent[word + 1] = 0;
    ent[word + 2] = (unsigned)b;
    ent[word + 3] = (unsigned)(b >> 32);
  }
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
  wave_release();
  ia = ja = nullptr;
  level_ptr.clear();
  n = 0;
}

void IluFactor::wave_release()
{
  for(auto &w : wave) {
    dfree(w.rk0);
    dfree(w.rcnt);
    dfree(w.row);
    dfree(w.aux);
    dfree(w.meta);
    dfree(w.ent);
    dfree(w.dinv32);
    dfree(w.dinv64);
    dfree(w.pre64);
    dfree(w.pre32);
    w = WaveSweep();
  }
  dfree(wave_ecol);
  wave_mode = false;
}

// The one-wave sweep streams for the host pattern (lower rows lrows grouped by
// level through lptr).  They are built when every row's entries fit one of the
// kernel's (lanes, slots) shapes, every dependency is still in the LDS ring when read,
// and the estimated time per application beats the sync-free sweeps': about
// 2 us per dependency level there (one publication round trip) against about
// 0.25 us per chunk here.  GDP_ILU_SWEEP=wave / free forces either choice; a
// forced `wave` on a system the kernel cannot take is GDP_E_UNSUPPORTED.
int IluFactor::setup_wave(const int *hia, const int *hja, const std::vector<int> &lrows,
                          const std::vector<int> &lptr)
{
  wave_release();
  const char *mode = std::getenv("GDP_ILU_SWEEP");
  // forced: a system the kernel cannot take is an error, not a silent fallback
  const bool force = mode && !std::strcmp(mode, "wave");
  // GDP_ILU_LANES selects a lane count of the sync-free sweeps
  if(!n || (mode && !std::strcmp(mode, "free")) || (!force && std::getenv("GDP_ILU_LANES"))) return GDP_OK;
  std::vector<int> dpos(n);
  for(int i = 0; i < n; i++) {
    int k = hia[i];
    while(k < hia[i + 1] && hja[k] < i) k++;
    if(k == hia[i + 1] || hja[k] != i)  // missing diagonal: reported by the caller
      return force ? set_error(GDP_E_UNSUPPORTED, "one-wave sweeps: row %d has no diagonal entry", i) : GDP_OK;
    dpos[i] = k;
  }
  std::vector<int> ulev;
  const int nul = upper_levels(n, hia, hja, ulev);
  std::vector<int> uptr(nul + 1, 0), urows(n);
  for(int i = 0; i < n; i++) uptr[ulev[i] + 1]++;
  for(int l = 0; l < nul; l++) uptr[l + 1] += uptr[l];
  {
    std::vector<int> f(uptr.begin(), uptr.end() - 1);
    for(int i = n - 1; i >= 0; i--) urows[f[ulev[i]]++] = i;
  }
  // entries of row i on side s: [b, e)
  auto range = [&](int s, int i, int &b, int &e) {
    if(s == 0) { b = hia[i]; e = dpos[i]; }
    else { b = dpos[i] + 1; e = hia[i + 1]; }
  };
  struct Side {
    int G, wg, rc;
    std::vector<int> cstart, ccount;  // chunk -> rows (index into the ordered rows)
    std::vector<int> spos;            // row -> stream position: its index in level order (ring slot mod kRing)
  } sd[2];
  const std::vector<int> *rows[2] = {&lrows, &urows};
  const std::vector<int> *ptrs[2] = {&lptr, &uptr};
  for(int s = 0; s < 2; s++) {
    int maxc = 0;
    for(int i = 0; i < n; i++) {
      int b, e;
      range(s, i, b, e);
      maxc = std::max(maxc, e - b);
    }
    int shape = 0;
    while(shape < kNumWaveShapes && kWaveSlots[shape] < maxc) shape++;
    if(shape == kNumWaveShapes)
      return force ? set_error(GDP_E_UNSUPPORTED, "one-wave sweeps: a row holds %d entries on one side (at most %d)",
                               maxc, kWaveSlots[kNumWaveShapes - 1])
                   : GDP_OK;
    sd[s].G = 1;
    sd[s].wg = kWaveSlots[shape];
    sd[s].rc = kWave / sd[s].G;
  }
  // Estimated time per application, before the chunk and ring analysis (a
  // pass over every entry), which a system the walk would not win skips.  A
  // chunk of the walk costs its reads and its row chain (about 0.25 us with
  // 4 entry slots a row, 0.46 us with 24), and the chunks of one level go
  // about two at a time; the sync-free sweeps cost 1-2 us per level (one
  // publication round trip).
  {
    double t_wave = 0.;
    for(int s = 0; s < 2; s++) {
      const int cap = sd[s].G * sd[s].wg;
      const double per = cap <= 4 ? 0.25 : cap <= 8 ? 0.3 : cap <= 16 ? 0.4 : cap <= 24 ? 0.46 : 0.55;
      const std::vector<int> &p = *ptrs[s];
      for(size_t l = 0; l + 1 < p.size(); l++) {
        const int ch = (p[l + 1] - p[l] + sd[s].rc - 1) / sd[s].rc;
        t_wave += per * std::max(1, (ch + 1) / 2);
      }
    }
    const double t_free = 2.0 * ((double)nlevels + upper_nlevels);
    if(!force && t_wave >= t_free) return GDP_OK;
  }
  for(int s = 0; s < 2; s++) {
    const std::vector<int> &p = *ptrs[s];
    for(size_t l = 0; l + 1 < p.size(); l++)
      for(int a = p[l]; a < p[l + 1]; a += sd[s].rc) {
        sd[s].cstart.push_back(a);
        sd[s].ccount.push_back(std::min(sd[s].rc, p[l + 1] - a));
      }
    // the ring holds the solved entries of the last kRing stream positions
    // (a row's position is its index in level order, its slot the position
    // mod kRing, so the ring empties in solve order and padding lanes take
    // no slots); every dependency must still be there when its chunk reads
    // it, i.e. solved fewer than kRing positions before the chunk starts
    const std::vector<int> &ord = *rows[s];
    sd[s].spos.assign(n, -1);
    for(size_t c = 0; c < sd[s].cstart.size(); c++)
      for(int t = 0; t < sd[s].ccount[c]; t++) sd[s].spos[ord[sd[s].cstart[c] + t]] = sd[s].cstart[c] + t;
    for(size_t c = 0; c < sd[s].cstart.size(); c++) {
      for(int t = 0; t < sd[s].ccount[c]; t++) {
        const int i = ord[sd[s].cstart[c] + t];
        int b, e;
        range(s, i, b, e);
        for(int k = b; k < e; k++)
          if((long)sd[s].cstart[c] - sd[s].spos[hja[k]] >= kRing - kWave)
            return force ? set_error(GDP_E_UNSUPPORTED,
                                     "one-wave sweeps: row %d depends on row %d, which has left the %d-entry ring", i,
                                     hja[k], kRing)
                         : GDP_OK;
      }
    }
  }
  // waves per walk: four (one per SIMD), which share the instruction issue of
  // the chunks' reads and products; GDP_ILU_TEAM overrides.  Measured per
  // application (1 and 4 waves): ILUTP(20) of Poisson 512^2 (one chunk a
  // level) 18.3 and 17.5 ms, 256^2 10.0 and 9.5 ms; ILU(0) of Poisson 256^2
  // in RCMK order (two chunks a level) 0.45 and 0.40 ms.
  int team = 4;
  if(const char *e = std::getenv("GDP_ILU_TEAM")) team = std::max(1, std::min(kMaxTeam, atoi(e)));
  // need[c] for the team walk (sequential chunk order, in which the ring check
  // above holds): the chunks of c's dependencies, and for every slot c writes
  // its previous writer and the last reader of the value it holds
  std::vector<int> need[2];
  for(int s = 0; s < 2; s++) {
    const Side &S = sd[s];
    const std::vector<int> &ord = *rows[s];
    const long nch = (long)S.cstart.size();
    std::vector<int> chunk_of(n, -1), writer(kRing, -1), reader(kRing, -1);
    need[s].assign(nch, 0);
    for(long c = 0; c < nch; c++) {
      int nd = 0;
      for(int t = 0; t < S.ccount[c]; t++) {
        const int i = ord[S.cstart[c] + t];
        int b, e;
        range(s, i, b, e);
        for(int k = b; k < e; k++) nd = std::max(nd, chunk_of[hja[k]] + 1);
        const int sl = S.spos[i] & (kRing - 1);
        nd = std::max(nd, std::max(writer[sl], reader[sl]) + 1);
      }
      for(int t = 0; t < S.ccount[c]; t++) {
        const int i = ord[S.cstart[c] + t];
        int b, e;
        range(s, i, b, e);
        for(int k = b; k < e; k++) reader[S.spos[hja[k]] & (kRing - 1)] = (int)c;
      }
      for(int t = 0; t < S.ccount[c]; t++) {
        const int i = ord[S.cstart[c] + t], sl = S.spos[i] & (kRing - 1);
        writer[sl] = (int)c;
        reader[sl] = -1;
        chunk_of[i] = (int)c;
      }
      need[s][c] = nd;
    }
  }
  // positions of the rows in the backward stream (the forward sweep's output)
  std::vector<int> bpos(n);
  for(size_t c = 0; c < sd[1].cstart.size(); c++)
    for(int t = 0; t < sd[1].ccount[c]; t++) bpos[urows[sd[1].cstart[c] + t]] = (int)c * kWave + t * sd[1].G;
  std::vector<int> ecol(std::max<long>(1, (long)hia[n]), kRing);
  for(int s = 0; s < 2; s++) {
    const Side &S = sd[s];
    const long nch = (long)S.cstart.size();
    // steps rounded up to whole register rotations of every wave, then the
    // padding chunks the last prefetches read
    const long span = (long)team * kWalkPad;
    const long nsteps = (nch + span - 1) / span * span, nalloc = nsteps + span;
    const size_t ne = (size_t)nalloc * S.wg * kWave, nl = (size_t)nalloc * kWave;
    std::vector<int> rk0(nl, 0), rcnt(nl, 0), row(nl, -1), aux(nl, -1);
    std::vector<int4> meta(nl, make_int4(-1, -1, kRing, 0));
    const std::vector<int> &ord = *rows[s];
    for(long c = 0; c < nch; c++)
      for(int t = 0; t < S.ccount[c]; t++) {
        const int i = ord[S.cstart[c] + t], gl = t * S.G;
        int b, e;
        range(s, i, b, e);
        // the row's entries are b .. e - 1 of the factors, and entry k's
        // dependency sits in ring slot ecol[k]
        for(int k = b; k < e; k++) ecol[k] = S.spos[hja[k]] & (kRing - 1);
        rk0[(size_t)c * kWave + gl] = b;
        rcnt[(size_t)c * kWave + gl] = e - b;
        row[(size_t)c * kWave + gl] = i;
        aux[(size_t)c * kWave + gl] = s == 0 ? bpos[i] : dpos[i];
        meta[(size_t)c * kWave + gl] = make_int4(i, s == 0 ? bpos[i] : -1, S.spos[i] & (kRing - 1), 0);
      }
    for(long c = 0; c < nch; c++)
      for(int l = 0; l < kWave; l++) meta[(size_t)c * kWave + l].w = team > 1 ? need[s][c] : 0;
    WaveSweep &W = wave[s];
    W.G = S.G;
    W.wg = S.wg;
    W.nch = (int)nsteps;
    W.team = team;
    GDP_TRY(dalloc(&W.rk0, nl));
    GDP_TRY(dalloc(&W.rcnt, nl));
    GDP_TRY(dalloc(&W.row, nl));
    GDP_TRY(dalloc(&W.aux, nl));
    GDP_TRY(dalloc(&W.meta, nl));
    GDP_TRY(dalloc(&W.pre64, nl));
    GDP_TRY(dalloc(&W.pre32, nl));
    GDP_TRY(h2d(W.rk0, rk0.data(), nl));
    GDP_TRY(h2d(W.rcnt, rcnt.data(), nl));
    GDP_TRY(h2d(W.row, row.data(), nl));
    GDP_TRY(h2d(W.aux, aux.data(), nl));
    GDP_TRY(h2d(W.meta, meta.data(), nl));
  }
  GDP_TRY(dalloc(&wave_ecol, ecol.size()));
  GDP_TRY(h2d(wave_ecol, ecol.data(), ecol.size()));
  wave_mode = true;
  return GDP_OK;
}

// Chunks stored per stream: the walk's chunks and one prefetch round per wave.
long IluFactor::wave_chunks_alloc(const WaveSweep &W) const { return (long)W.nch + (long)W.team * kWalkPad; }

// Storage for the packed factor entries, in the factors' precision.
int IluFactor::wave_alloc_values()
{
  if(!wave_mode) return GDP_OK;
  for(int s = 0; s < 2; s++) {
    WaveSweep &W = wave[s];
    const size_t nalloc = (size_t)wave_chunks_alloc(W), nl = nalloc * kWave;
    dfree(W.ent);
    dfree(W.dinv32);
    dfree(W.dinv64);
    GDP_TRY(dalloc(&W.ent, nl * walk_nl(is_f64, W.wg) * 4));  // 16-byte records
    if(s == 1) {
      if(is_f64) GDP_TRY(dalloc(&W.dinv64, nl));
      else GDP_TRY(dalloc(&W.dinv32, nl));
    }
  }
  return GDP_OK;
}

// Copy the factor entries into the streams (enqueued: usable under capture).
void IluFactor::wave_pack()
{
  if(!wave_mode) return;
  hipStream_t st = ctx().stream;
  for(int s = 0; s < 2; s++) {
    WaveSweep &W = wave[s];
    const long nalloc = wave_chunks_alloc(W), ne = nalloc * W.wg * kWave, nl = nalloc * kWave;
    const dim3 ge((unsigned)((ne + 255) / 256)), gl((unsigned)((nl + 255) / 256));
    if(is_f64) {
      hipLaunchKernelGGL(k_wave_pack_ent<double>, ge, dim3(256), 0, st, ne, W.wg, W.rk0, W.rcnt, wave_ecol, lu64,
                         W.ent);
      if(s == 1) hipLaunchKernelGGL(k_wave_pack<double>, gl, dim3(256), 0, st, nl, W.aux, lu64, W.dinv64);
    }
    else {
      hipLaunchKernelGGL(k_wave_pack_ent<float>, ge, dim3(256), 0, st, ne, W.wg, W.rk0, W.rcnt, wave_ecol, lu32,
                         W.ent);
      if(s == 1) hipLaunchKernelGGL(k_wave_pack<float>, gl, dim3(256), 0, st, nl, W.aux, lu32, W.dinv32);
    }
  }
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
  return setup_wave(hia, hja, rows, level_ptr);
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
  GDP_TRY(wave_alloc_values());
  wave_pack();
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
  GDP_TRY(wave_alloc_values());
  if(f64) enqueue_levels(lu64);
  else enqueue_levels(lu32);
  wave_pack();
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
  return wave_alloc_values();
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
  wave_pack();
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

template <typename TF> const TF *wave_dinv(const IluFactor::WaveSweep &W);
template <> const float *wave_dinv<float>(const IluFactor::WaveSweep &W) { return W.dinv32; }
template <> const double *wave_dinv<double>(const IluFactor::WaveSweep &W) { return W.dinv64; }
template <typename TV> TV *wave_pre(const IluFactor::WaveSweep &W);
template <> float *wave_pre<float>(const IluFactor::WaveSweep &W) { return W.pre32; }
template <> double *wave_pre<double>(const IluFactor::WaveSweep &W) { return W.pre64; }

// x = (LU)^{-1} (scale * r) by the one-wave sweeps: gather scale * r into the
// forward stream's order, forward sweep (its output in the backward stream's
// order), backward sweep into x (and addto).
template <typename TF, typename TV>
void wave_apply(int n, const IluFactor::WaveSweep *W, const TV *r, TV scale, TV *x, TV *addto, int *err,
                unsigned long long tl, const int *gate, int gmode)
{
  hipStream_t s = ctx().stream;
  const IluFactor::WaveSweep &F = W[0], &B = W[1];
  const long nl = ((long)F.nch + (long)F.team * kWalkPad) * kWave;
  hipLaunchKernelGGL(k_wave_gather<TV>, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, s, nl, F.row, r, scale,
                     wave_pre<TV>(F), gate, gmode);
#define GDP_WAVE(GG, WW, FWD, S, OUT, ADD)                                                                \
  hipLaunchKernelGGL((k_wave_sweep<TF, TV, WW, FWD>), dim3(1), dim3(kWave * S.team), 0, s, S.nch, \
                     (const uint4 *)S.ent, S.meta, wave_dinv<TF>(S), wave_pre<TV>(S), OUT, err, tl, gate, gmode)
#define GDP_WAVE_SHAPE(FWD, S, OUT, ADD)                                        \
  if(S.wg == 4) GDP_WAVE(1, 4, FWD, S, OUT, ADD);                               \
  else if(S.wg == 8) GDP_WAVE(1, 8, FWD, S, OUT, ADD);                          \
  else if(S.wg == 16) GDP_WAVE(1, 16, FWD, S, OUT, ADD);                        \
  else if(S.wg == 24) GDP_WAVE(1, 24, FWD, S, OUT, ADD);                        \
  else GDP_WAVE(1, 32, FWD, S, OUT, ADD)
  GDP_WAVE_SHAPE(true, F, wave_pre<TV>(B), (TV *)nullptr);
  GDP_WAVE_SHAPE(false, B, x, addto);
#undef GDP_WAVE_SHAPE
#undef GDP_WAVE
  if(addto) hipLaunchKernelGGL(k_wave_addto<TV>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, addto, gate, gmode);
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
  if(wave_mode) {
    if(is_f64) wave_apply<double, TV>(n, wave, r, scale, out, add, err + 1, wait_limit(), gate, gmode);
    else wave_apply<float, TV>(n, wave, r, scale, out, add, err + 1, wait_limit(), gate, gmode);
  }
  else if(is_f64) {
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
