// Batched ILU(0)-GMRES over many independent systems.
//
// The systems are the diagonal blocks of one matrix (their meshes are
// concatenated into one formulation, so the DOFs of instance i are the rows
// [off[i], off[i+1])).  Per instance the method is the reference's
// solve_matrix with Algorithm 8 / Preconditioner 6 (Sparskit.cpp:108-634):
// renumbering, ILU(0) with FP32 factor storage, right-preconditioned
// GMRES(m) from x0 = 0 stopping on the estimated residual against |r0|.
//
// Execution: one persistent kernel per launch configuration.  A workgroup
// takes the next instance from a queue ordered by decreasing size and runs
// the whole solve on it: SpMV over its rows, the triangular sweeps level by
// level with a workgroup barrier between levels, Gram-Schmidt with
// workgroup reductions, Givens rotations in LDS.  Instances never
// communicate, so no grid-wide synchronisation exists.  Assembly, value
// permutation, factorisation, the solve kernels and the final unpermutation
// are captured once into a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <type_traits>

#include "gdp_device.h"
#include "ilu.h"

namespace gdp {

int fe_generate_enqueue(gdp_fe *fe, gdp_matrix *A, gdp_vector *b);

namespace {

constexpr int kBM = 32;  // largest Krylov size of the batched kernel

struct BatchArgs {
  int m, maxits, orth;
  double rtol;
  const int *ia, *ja;
  const double *a;
  const int *diag;
  const float *lu;
  const int *off;
  // sweep levels per instance: slots ptr[base + l] .. ptr[base + l + 1] of
  // level l; a slot is one row's record (local row, entry count, local
  // columns and FP32 factor values padded to W, inverted pivot)
  const int *fl_base, *fl_n, *fl_ptr;
  const int *bl_base, *bl_n, *bl_ptr;
  const int *f_row, *f_cnt, *f_col;
  const float *f_val;
  const int *b_row, *b_cnt, *b_col;
  const float *b_val, *b_dinv;
  // per slot {row, entry count, level within the instance, 0}: the fields the
  // 1024-thread walk streams into LDS
  const int4 *f_meta, *b_meta;
  int zcap;  // instances with at most zcap rows keep the sweep vector in LDS
  // larger instances with ring[inst] keep the last rmask + 1 sweep results in
  // an LDS ring indexed by stream position (their record columns are stream
  // positions); the others read the sweep vector from global memory
  const int *ring;
  int rmask;
  int walk;  // sweeps of instances with the sweep vector in LDS walked by one wave (else level by level)
  const long *pool_off;
  double *pool;
  const double *rhs;
  double *sol;
  const int *order;  // instances of this launch, largest first
  int count;
  int *queue;
  int *its, *status;
  double *res0, *res;
  unsigned long long *t0, *t1;
};

template <int T> __device__ inline double block_sum_d(double v, double *red)
{
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if(l == 0) red[w] = v;
  __syncthreads();
  double r = 0.;
#pragma unroll
  for(int i = 0; i < T / 64; i++) r += red[i];
  return r;  // every thread holds the (identically ordered) sum
}

// y = A x over the instance's rows (global row indices)
__device__ inline void b_spmv(const BatchArgs &a, int r0, int n, const double *x, double *y, int T)
{
  for(int i = threadIdx.x; i < n; i += T) {
    const int row = r0 + i;
    double s = 0.;
    for(int k = a.ia[row]; k < a.ia[row + 1]; k++) s += a.a[k] * x[a.ja[k] - r0];
    y[i] = s;
  }
}

// One row of a sweep: what a thread needs besides the solution entries it
// depends on (loaded a level ahead, off the dependency chain).
template <int W> struct SweepRec {
  int row, cnt, pos;  // pos: the record's position in the instance's sweep stream
  int col[W];
  float val[W];
  float dinv;
  double pre;  // forward: v[row]; backward: the forward result z[row]
};

template <int W, bool FWD>
__device__ inline void ld_rec(const BatchArgs &a, int slot, int s_first, const double *src, SweepRec<W> &r)
{
  const int *rowp = FWD ? a.f_row : a.b_row, *cntp = FWD ? a.f_cnt : a.b_cnt;
  const int *colp = FWD ? a.f_col : a.b_col;
  const float *valp = FWD ? a.f_val : a.b_val;
  r.row = rowp[slot];
  r.cnt = cntp[slot];
  r.pos = slot - s_first;
#pragma unroll
  for(int w = 0; w < W; w++) {
    r.col[w] = colp[(long)slot * W + w];
    r.val[w] = valp[(long)slot * W + w];
  }
  r.dinv = FWD ? 1.f : a.b_dinv[slot];
  r.pre = src[r.row];
}

// The same record with its right-hand side from P, the sweep's inputs staged
// in slot order (no gather behind the load of the row).
template <int W, bool FWD>
__device__ inline void ld_rec_staged(const BatchArgs &a, int slot, int s_first, const double *P, SweepRec<W> &r)
{
  const int *rowp = FWD ? a.f_row : a.b_row, *cntp = FWD ? a.f_cnt : a.b_cnt;
  const int *colp = FWD ? a.f_col : a.b_col;
  const float *valp = FWD ? a.f_val : a.b_val;
  r.row = rowp[slot];
  r.cnt = cntp[slot];
  r.pos = slot - s_first;
#pragma unroll
  for(int w = 0; w < W; w++) {
    r.col[w] = colp[(long)slot * W + w];
    r.val[w] = valp[(long)slot * W + w];
  }
  r.dinv = FWD ? 1.f : a.b_dinv[slot];
  r.pre = P[slot - s_first];
}

// P[q] = src[row of slot q] over an instance's slots of one sweep (all threads)
template <int T, bool FWD>
__device__ inline void stage_inputs(const BatchArgs &a, int inst, const double *src, double *P)
{
  const int *ptr = FWD ? a.fl_ptr : a.bl_ptr;
  const int base = FWD ? a.fl_base[inst] : a.bl_base[inst];
  const int nl = FWD ? a.fl_n[inst] : a.bl_n[inst];
  if(nl <= 0) return;
  const int *rowp = FWD ? a.f_row : a.b_row;
  const int sf = ptr[base], se = ptr[base + nl];
  // eight slots a thread at a time: their row loads, then their gathers
  for(int q0 = sf + (int)threadIdx.x; q0 < se; q0 += 8 * T) {
    int row[8];
#pragma unroll
    for(int u = 0; u < 8; u++) row[u] = q0 + u * T < se ? rowp[q0 + u * T] : -1;
    double v[8];
#pragma unroll
    for(int u = 0; u < 8; u++) v[u] = row[u] >= 0 ? src[row[u]] : 0.;
#pragma unroll
    for(int u = 0; u < 8; u++)
      if(row[u] >= 0) P[q0 + u * T - sf] = v[u];
  }
}

// lusol (ilut.F:2297-2347) on one row, same operation order:
// forward x(i) = y(i) - sum alu(k) x(jlu(k)); backward x(i) = alu(i) (x(i) - sum ...)
// RING: the solution entries come from the LDS ring (column = stream
// position), and the result goes to both the ring and z
template <int W, bool FWD, bool RING>
__device__ inline void apply_rec(const SweepRec<W> &r, double *z, double *ring, int rmask)
{
  // every dependency is read before the first subtraction (padding columns
  // are 0, in range, and their values are not used): one read latency per
  // row instead of one per entry
  double dv[W];
#pragma unroll
  for(int w = 0; w < W; w++) dv[w] = RING ? ring[r.col[w] & rmask] : z[r.col[w]];
  double acc = r.pre;
#pragma unroll
  for(int w = 0; w < W; w++)
    if(w < r.cnt) acc = acc - (double)r.val[w] * dv[w];
  const double x = FWD ? acc : (double)r.dinv * acc;
  z[r.row] = x;
  if constexpr(RING) ring[r.pos & rmask] = x;
}

// Barrier between sweep levels.  With the sweep vector in LDS only LDS
// accesses need ordering, so the fence is restricted to the local address
// space: the next level's record loads (global memory) stay in flight across
// the barrier instead of being waited for.
template <bool LDSZ> __device__ inline void level_barrier()
{
  if constexpr(LDSZ) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
  else
    __syncthreads();
}

// One triangular sweep, level by level with a barrier between levels; the
// next level's records are loaded before the current level is computed.
template <int T, int W, bool FWD, bool LDSZ, bool RING = false>
__device__ inline void sweep(const BatchArgs &a, int inst, const double *src, double *z, double *ring = nullptr)
{
  // software pipeline over levels: a level's records are loaded kAhead levels
  // before it is computed, and its slot range kAhead levels before that, so
  // that no global load sits on the level-to-level chain
  constexpr int kAhead = 3;
  const int *ptr = FWD ? a.fl_ptr : a.bl_ptr;
  const int base = FWD ? a.fl_base[inst] : a.bl_base[inst];
  const int nl = FWD ? a.fl_n[inst] : a.bl_n[inst];
  if(nl <= 0) return;
  const int tid = threadIdx.x;
  const int sf = ptr[base];
  SweepRec<W> buf[kAhead];
  int bs0[kAhead], bs1[kAhead], ns0[kAhead], ns1[kAhead];
  bool bh[kAhead];
#pragma unroll
  for(int d = 0; d < kAhead; d++) {
    bs0[d] = d < nl ? ptr[base + d] : 0;
    bs1[d] = d < nl ? ptr[base + d + 1] : 0;
    ns0[d] = d + kAhead < nl ? ptr[base + d + kAhead] : 0;
    ns1[d] = d + kAhead < nl ? ptr[base + d + kAhead + 1] : 0;
  }
#pragma unroll
  for(int d = 0; d < kAhead; d++) {
    bh[d] = bs0[d] + tid < bs1[d];
    if(bh[d]) ld_rec<W, FWD>(a, bs0[d] + tid, sf, src, buf[d]);
  }
  for(int l0 = 0; l0 < nl; l0 += kAhead) {
#pragma unroll
    for(int d = 0; d < kAhead; d++) {
      const int l = l0 + d;
      if(l >= nl) break;
      const SweepRec<W> cur = buf[d];
      const bool has = bh[d];
      const int s0 = bs0[d], s1 = bs1[d];
      // level l + kAhead: its records; level l + 2 kAhead: its slot range
      bs0[d] = ns0[d];
      bs1[d] = ns1[d];
      bh[d] = l + kAhead < nl && bs0[d] + tid < bs1[d];
      if(bh[d]) ld_rec<W, FWD>(a, bs0[d] + tid, sf, src, buf[d]);
      const int l2 = l + 2 * kAhead;
      ns0[d] = l2 < nl ? ptr[base + l2] : 0;
      ns1[d] = l2 < nl ? ptr[base + l2 + 1] : 0;
      if(has) apply_rec<W, FWD, RING>(cur, z, ring, a.rmask);
      for(int q = s0 + tid + T; q < s1; q += T) {
        SweepRec<W> r;
        ld_rec<W, FWD>(a, q, sf, src, r);
        apply_rec<W, FWD, RING>(r, z, ring, a.rmask);
      }
      level_barrier<LDSZ>();
    }
  }
}

// One triangular sweep walked by one wave (the caller's wave 0; the other
// waves wait at the caller's barrier).  With the sweep vector in global
// memory (instances that read their dependencies from the LDS ring), the
// inputs are first staged in slot order (STAGED), so that a record's
// right-hand side is not a gather behind the load of its row: three windows
// ahead then cover one memory round trip instead of two.  The wave takes the instance's slots 64
// at a time in stream (level) order; inside such a window it solves the rows
// of one level after the other, so a row that depends on a row of the same
// window finds it solved: the wave's LDS accesses take effect in program
// order, and every dependency is in LDS (the sweep vector, or the ring).  No
// barrier separates the levels; the records of the next D windows are in
// flight while a window is solved.  Same operations per row as apply_rec.
template <int W, bool FWD, bool RING, bool STAGED = false>
__device__ inline void walk_sweep(const BatchArgs &a, int inst, const double *src, double *z, double *ring)
{
  auto load = [&](int q, int sf, SweepRec<W> &r) {
    if constexpr(STAGED) ld_rec_staged<W, FWD>(a, q, sf, src, r);
    else ld_rec<W, FWD>(a, q, sf, src, r);
  };
  constexpr int D = 3;
  const int *ptr = FWD ? a.fl_ptr : a.bl_ptr;
  const int base = FWD ? a.fl_base[inst] : a.bl_base[inst];
  const int nl = FWD ? a.fl_n[inst] : a.bl_n[inst];
  if(nl <= 0) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int sf = ptr[base], se = ptr[base + nl];
  SweepRec<W> buf[D];
  bool has[D];
  // loads are unconditional (a lane past the end reloads the last slot):
  // a load under a branch made the compiler wait for every load in flight
  // where the branches join, once per window
#pragma unroll
  for(int d = 0; d < D; d++) {
    const int q = sf + kWave * d + lane;
    has[d] = q < se;
    load(has[d] ? q : se - 1, sf, buf[d]);
  }
  // Level ends: lane j holds the end slot of level lb + j, and of level
  // lb + 64 + j in the next block, loaded half a block ahead; a level's end
  // is read from its lane (no memory access between two levels).
  int l = 0, lb = 0;
  int lcur = lane < nl ? ptr[base + 1 + lane] : 0x7fffffff;
  int lnext = lane + kWave < nl ? ptr[base + 1 + kWave + lane] : 0x7fffffff;
  auto level_end = [&](int lv) {
    if(lv - lb >= kWave) {
      lb += kWave;
      lcur = lnext;
      lnext = lb + kWave + lane < nl ? ptr[base + 1 + lb + kWave + lane] : 0x7fffffff;
    }
    return __builtin_amdgcn_readlane(lcur, lv - lb);
  };
  int le = level_end(0);  // the current level's end slot
  for(int w0 = sf; w0 < se; w0 += kWave * D) {
#pragma unroll
    for(int d = 0; d < D; d++) {
      const int w = w0 + kWave * d;
      if(w >= se) break;
      const SweepRec<W> cur = buf[d];
      bool done = !has[d];
      {
        const int q = w + kWave * D + lane;  // this buffer's next window
        has[d] = q < se;
        load(has[d] ? q : se - 1, sf, buf[d]);
      }
      const int wend = min(w + kWave, se), slot = w + lane;
      while(le <= w) le = level_end(++l);
      for(;;) {
        if(!done && slot < le) {
          apply_rec<W, FWD, RING>(cur, z, ring, a.rmask);
          done = true;
        }
        if(le >= wend) break;
        le = level_end(++l);
      }
    }
  }
}

// ---- the 1024-thread walk: records streamed into LDS --------------------------
//
// The walk above keeps D = 3 windows of records in registers, and at about a
// window per 0.7 us it waits on their loads: three windows of work do not
// cover a memory round trip, and deeper register buffers do not fit the
// 1024-thread kernel.  Here the walking wave streams the records of the next
// F - 1 windows straight into an LDS FIFO with LDS-DMA loads
// (global_load_lds: no registers held while in flight), waits for the oldest
// window with a vector-memory count (loads complete in order, so with
// (F - 1) * L loads issued after a window's L loads, vmcnt <= (F - 1) * L means
// the window has landed), and reads it with LDS reads.  The FIFO is its own
// __shared__ object: the compiler then knows the sweep vector and the ring
// cannot alias it and puts no wait for the DMA before their accesses.  The
// FIFO reads are inline asm for the same reason (the compiler would wait for
// every DMA in flight before a read of the FIFO); they wait for themselves in
// the same asm block.  Same operations per row as apply_rec; levels come
// from each slot's record, so the walk touches no other memory.
template <int W> struct WalkFifo {
  static constexpr int F = W == 4 ? 8 : 4;  // windows in the FIFO
  int4 meta[F][kWave];
  int4 col[F][W / 4][kWave];
  uint4 val[F][W / 4][kWave];
  unsigned dinv[F][kWave];
  unsigned plo[F][kWave], phi[F][kWave];
};

#define GDP_LDS __attribute__((address_space(3)))
__device__ inline void dma16(const void *g, void *l)
{
  __builtin_amdgcn_global_load_lds(g, (GDP_LDS void *)l, 16, 0, 0);
}
__device__ inline void dma4(const void *g, void *l) { __builtin_amdgcn_global_load_lds(g, (GDP_LDS void *)l, 4, 0, 0); }
__device__ inline unsigned lds_addr(const void *p) { return (unsigned)(size_t)(const GDP_LDS void *)p; }
// s_waitcnt vmcnt(N) (gfx9 encoding: vmcnt bits 3:0 and 15:14), other counters untouched
template <int N> __device__ inline void wait_vmcnt()
{
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <int W, bool FWD, bool RING>
__device__ inline void walk_fifo(const BatchArgs &a, int inst, const double *P, double *z, double *ring,
                                 WalkFifo<W> &fifo)
{
  constexpr int F = WalkFifo<W>::F;
  constexpr int L = 1 + 2 * (W / 4) + (FWD ? 0 : 1) + 2;  // DMA loads per window
  const int *ptr = FWD ? a.fl_ptr : a.bl_ptr;
  const int base = FWD ? a.fl_base[inst] : a.bl_base[inst];
  const int nl = FWD ? a.fl_n[inst] : a.bl_n[inst];
  if(nl <= 0) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int sf = ptr[base], se = ptr[base + nl];
  const int nwin = (se - sf + kWave - 1) / kWave;
  const int4 *metap = FWD ? a.f_meta : a.b_meta;
  const int *colp = FWD ? a.f_col : a.b_col;
  const float *valp = FWD ? a.f_val : a.b_val;
  const unsigned *pw = (const unsigned *)P;
  // window k into FIFO slot k % F (past the end: the last window again, so
  // that every step issues the same L loads; lanes past the end: the last slot)
  auto issue = [&](int k) {
    const int fs = k & (F - 1);
    const int q = min(sf + min(k, nwin - 1) * kWave + lane, se - 1);
    dma16(metap + q, &fifo.meta[fs][0]);
#pragma unroll
    for(int j = 0; j < W / 4; j++) dma16(colp + (size_t)q * W + 4 * j, &fifo.col[fs][j][0]);
#pragma unroll
    for(int j = 0; j < W / 4; j++) dma16(valp + (size_t)q * W + 4 * j, &fifo.val[fs][j][0]);
    if constexpr(!FWD) dma4(a.b_dinv + q, &fifo.dinv[fs][0]);
    dma4(pw + 2 * (size_t)(q - sf), &fifo.plo[fs][0]);
    dma4(pw + 2 * (size_t)(q - sf) + 1, &fifo.phi[fs][0]);
  };
#pragma unroll
  for(int d = 0; d + 1 < F; d++) issue(d);
  for(int k = 0; k < nwin; k++) {
    issue(k + F - 1);
    wait_vmcnt<(F - 1) * L>();
    const int fs = k & (F - 1);
    int4 m;
    int cl[W];
    unsigned vl[W], di, lo, hi;
    if constexpr(W == 4) {
      int4 c;
      uint4 v;
      asm volatile("ds_read_b128 %0, %6\n\t"
                   "ds_read_b128 %1, %7\n\t"
                   "ds_read_b128 %2, %8\n\t"
                   "ds_read_b32 %3, %9\n\t"
                   "ds_read_b32 %4, %10\n\t"
                   "ds_read_b32 %5, %11\n\t"
                   "s_waitcnt lgkmcnt(0)"
                   : "=&v"(m), "=&v"(c), "=&v"(v), "=&v"(di), "=&v"(lo), "=&v"(hi)
                   : "v"(lds_addr(&fifo.meta[fs][lane])), "v"(lds_addr(&fifo.col[fs][0][lane])),
                     "v"(lds_addr(&fifo.val[fs][0][lane])), "v"(lds_addr(&fifo.dinv[fs][lane])),
                     "v"(lds_addr(&fifo.plo[fs][lane])), "v"(lds_addr(&fifo.phi[fs][lane]))
                   : "memory");
      cl[0] = c.x, cl[1] = c.y, cl[2] = c.z, cl[3] = c.w;
      vl[0] = v.x, vl[1] = v.y, vl[2] = v.z, vl[3] = v.w;
    }
    else {
      int4 c0, c1;
      uint4 v0, v1;
      asm volatile("ds_read_b128 %0, %8\n\t"
                   "ds_read_b128 %1, %9\n\t"
                   "ds_read_b128 %2, %10\n\t"
                   "ds_read_b128 %3, %11\n\t"
                   "ds_read_b128 %4, %12\n\t"
                   "ds_read_b32 %5, %13\n\t"
                   "ds_read_b32 %6, %14\n\t"
                   "ds_read_b32 %7, %15\n\t"
                   "s_waitcnt lgkmcnt(0)"
                   : "=&v"(m), "=&v"(c0), "=&v"(c1), "=&v"(v0), "=&v"(v1), "=&v"(di), "=&v"(lo), "=&v"(hi)
                   : "v"(lds_addr(&fifo.meta[fs][lane])), "v"(lds_addr(&fifo.col[fs][0][lane])),
                     "v"(lds_addr(&fifo.col[fs][W / 4 - 1][lane])), "v"(lds_addr(&fifo.val[fs][0][lane])),
                     "v"(lds_addr(&fifo.val[fs][W / 4 - 1][lane])), "v"(lds_addr(&fifo.dinv[fs][lane])),
                     "v"(lds_addr(&fifo.plo[fs][lane])), "v"(lds_addr(&fifo.phi[fs][lane]))
                   : "memory");
      cl[0] = c0.x, cl[1] = c0.y, cl[2] = c0.z, cl[3] = c0.w;
      cl[W - 4] = c1.x, cl[W - 3] = c1.y, cl[W - 2] = c1.z, cl[W - 1] = c1.w;
      vl[0] = v0.x, vl[1] = v0.y, vl[2] = v0.z, vl[3] = v0.w;
      vl[W - 4] = v1.x, vl[W - 3] = v1.y, vl[W - 2] = v1.z, vl[W - 1] = v1.w;
    }
    const int slot = sf + k * kWave + lane;
    const int row = m.x, cnt = m.y, lvl = m.z;
    const double pre = __hiloint2double((int)hi, (int)lo);
    const bool valid = slot < se;
    bool done = !valid;
    double x = 0.;
    // the window's levels in order (levels are ascending in slot order): the
    // rows of one level, then the first lane still pending names the next
    int cur = __builtin_amdgcn_readfirstlane(lvl);
    for(;;) {
      if(!done && lvl == cur) {
        double dv[W];
#pragma unroll
        for(int w = 0; w < W; w++) dv[w] = RING ? ring[cl[w] & a.rmask] : z[cl[w]];
        double acc = pre;
#pragma unroll
        for(int w = 0; w < W; w++)
          if(w < cnt) acc = acc - (double)__uint_as_float(vl[w]) * dv[w];
        x = FWD ? acc : (double)__uint_as_float(di) * acc;
        if constexpr(RING) ring[(slot - sf) & a.rmask] = x;
        else z[row] = x;
        done = true;
      }
      const unsigned long long pend = __ballot(!done);
      if(!pend) break;
      cur = __builtin_amdgcn_readlane(lvl, __builtin_ctzll(pend));
    }
    // ring sweeps read their dependencies from the ring only: the results go
    // to the sweep vector in memory once per window (one store instruction)
    if constexpr(RING)
      if(valid) z[row] = x;
  }
  wait_vmcnt<0>();  // the padding windows' loads, before the FIFO is reused
}

// z = (LU)^{-1} v (v may alias z); z in LDS (zlds) when the instance fits
template <int T, int W, typename FIFO>
__device__ inline void b_ilu(const BatchArgs &a, int inst, int n, const double *v, double *z, double *zlds,
                             FIFO &fifo)
{
  if constexpr(T == 1024) {
    if(a.walk && (n <= a.zcap || a.ring[inst])) {
      // inputs staged in slot order in the instance's spare workspace vector,
      // then walked by wave 0 with the records streamed through the FIFO
      const bool lz = n <= a.zcap;
      double *P = a.pool + a.pool_off[inst] + (size_t)(a.m + 2) * n;
      stage_inputs<T, true>(a, inst, v, P);
      __syncthreads();
      if(threadIdx.x < kWave) {
        if(lz) walk_fifo<W, true, false>(a, inst, P, zlds, nullptr, fifo);
        else walk_fifo<W, true, true>(a, inst, P, z, zlds, fifo);
      }
      __syncthreads();
      stage_inputs<T, false>(a, inst, lz ? zlds : z, P);
      __syncthreads();
      if(threadIdx.x < kWave) {
        if(lz) walk_fifo<W, false, false>(a, inst, P, zlds, nullptr, fifo);
        else walk_fifo<W, false, true>(a, inst, P, z, zlds, fifo);
      }
      __syncthreads();
      return;
    }
  }
  if(n <= a.zcap && a.walk) {
    if(threadIdx.x < kWave) {
      walk_sweep<W, true, false>(a, inst, v, zlds, nullptr);
      walk_sweep<W, false, false>(a, inst, zlds, zlds, nullptr);
    }
    __syncthreads();
  }
  else if(a.ring[inst] && a.walk) {
    // inputs staged in slot order in the instance's spare workspace vector;
    // the forward results reach the backward inputs through z
    double *P = a.pool + a.pool_off[inst] + (size_t)(a.m + 2) * n;
    stage_inputs<T, true>(a, inst, v, P);
    __syncthreads();
    if(threadIdx.x < kWave) walk_sweep<W, true, true, true>(a, inst, P, z, zlds);
    __syncthreads();
    stage_inputs<T, false>(a, inst, z, P);
    __syncthreads();
    if(threadIdx.x < kWave) walk_sweep<W, false, true, true>(a, inst, P, z, zlds);
    __syncthreads();
  }
  else if(n <= a.zcap) {
    sweep<T, W, true, true>(a, inst, v, zlds);
    sweep<T, W, false, true>(a, inst, zlds, zlds);
    __syncthreads();  // the consumers of z also read global memory
  }
  else if(a.ring[inst]) {
    // dependencies through the LDS ring (LDS fences between levels); the
    // forward results reach the backward sweep's records through z
    sweep<T, W, true, true, true>(a, inst, v, z, zlds);
    __syncthreads();
    sweep<T, W, false, true, true>(a, inst, z, z, zlds);
    __syncthreads();
  }
  else {
    sweep<T, W, true, false>(a, inst, v, z);
    sweep<T, W, false, false>(a, inst, z, z);
  }
}

template <int T, int W, int ORTH>
__global__ void __launch_bounds__(T) k_batch_gmres(BatchArgs a)
{
  extern __shared__ double zlds[];
  // the record FIFO of the 1024-thread walk (a byte elsewhere)
  __shared__ std::conditional_t<T == 1024, WalkFifo<W>, char> fifo;
  __shared__ double H[(kBM + 1) * kBM];
  __shared__ double cs[kBM], sn[kBM], g[kBM + 1], yv[kBM];
  __shared__ double red[T / 64];
  __shared__ double mred[kBM + 1][T / 64];
  __shared__ int s_q;
  const int tid = threadIdx.x;
  const int m = a.m;
  const int maxits = a.maxits > 0 ? a.maxits : 1000000;
  for(;;) {
    __syncthreads();
    if(tid == 0) s_q = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = s_q;
    if(q >= a.count) break;
    const int inst = a.order[q];
    const int r0 = a.off[inst], n = a.off[inst + 1] - r0;
    if(tid == 0) a.t0[inst] = wall_clock64();
    double *V = a.pool + a.pool_off[inst];
    double *Z = n <= a.zcap ? zlds : V + (size_t)(m + 1) * n;
    double *x = a.sol + r0;
    const double *b = a.rhs + r0;
    // x0 = 0: r0 = b
    double bb = 0.;
    for(int i = tid; i < n; i += T) {
      x[i] = 0.;
      V[i] = b[i];
      bb += b[i] * b[i];
    }
    double beta = sqrt(block_sum_d<T>(bb, red));
    const double beta0 = beta, tol = a.rtol * beta;
    int its = 0, st = -1;
    double rest = beta;
    if(beta == 0.) st = 0;
    while(st < 0) {
      const double inv = 1. / beta;
      for(int i = tid; i < n; i += T) V[i] *= inv;
      if(tid == 0) {
        g[0] = beta;
        for(int i = 1; i <= m; i++) g[i] = 0.;
      }
      __syncthreads();
      int jend = -1;
      bool conv = false;
      for(int j = 0; j < m; j++) {
        double *vj = V + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
        b_ilu<T, W>(a, inst, n, vj, Z, zlds, fifo);
        b_spmv(a, r0, n, Z, w, T);
        its++;
        __syncthreads();
        if constexpr(ORTH == 1) {
          // classical Gram-Schmidt, twice
          for(int pass = 0; pass < 2; pass++) {
            // the j + 1 dot products 8 at a time (8 accumulators a thread
            // instead of 33: no spills at 1024 threads); each product sums
            // its rows in the same order as in one pass
            for(int i0 = 0; i0 <= j; i0 += 8) {
              double d[8];
#pragma unroll
              for(int u = 0; u < 8; u++) d[u] = 0.;
              for(int r = tid; r < n; r += T) {
                const double wr = w[r];
#pragma unroll
                for(int u = 0; u < 8; u++)
                  if(i0 + u <= j) d[u] += V[(size_t)(i0 + u) * n + r] * wr;
              }
#pragma unroll
              for(int u = 0; u < 8; u++)
                if(i0 + u <= j) {
                  const double s = wave_sum(d[u]);
                  if((tid & 63) == 0) mred[i0 + u][tid >> 6] = s;
                }
            }
            __syncthreads();
            if(tid <= j) {
              double s = 0.;
              for(int k = 0; k < T / 64; k++) s += mred[tid][k];
              if(pass == 0) H[tid * kBM + j] = s;
              else H[tid * kBM + j] += s;
              mred[tid][0] = s;  // coefficients of this pass
            }
            __syncthreads();
            for(int r = tid; r < n; r += T) {
              double wr = w[r];
              for(int i = 0; i <= j; i++) wr -= mred[i][0] * V[(size_t)i * n + r];
              w[r] = wr;
            }
            __syncthreads();
          }
        }
        else {
          // modified Gram-Schmidt; the update by v_i and the product with
          // v_{i+1} share one pass over w (a thread's rows are its own, so
          // the product reads w just updated, as after a separate pass)
          double h;
          {
            double dd = 0.;
            for(int r = tid; r < n; r += T) dd += V[r] * w[r];
            h = block_sum_d<T>(dd, red);
          }
          for(int i = 0; i <= j; i++) {
            const double *vi = V + (size_t)i * n;
            if(tid == 0) H[i * kBM + j] = h;
            if(i < j) {
              const double *vn = vi + n;
              double dd = 0.;
              for(int r = tid; r < n; r += T) {
                const double wr = w[r] - h * vi[r];
                w[r] = wr;
                dd += vn[r] * wr;
              }
              h = block_sum_d<T>(dd, red);
            }
            else
              for(int r = tid; r < n; r += T) w[r] -= h * vi[r];
          }
          __syncthreads();
        }
        double ww = 0.;
        for(int r = tid; r < n; r += T) ww += w[r] * w[r];
        const double hn = sqrt(block_sum_d<T>(ww, red));
        if(hn > 0.) {
          const double ih = 1. / hn;
          for(int r = tid; r < n; r += T) w[r] *= ih;
        }
        if(tid == 0) {
          H[(j + 1) * kBM + j] = hn;
          // apply the previous rotations to column j, then a new one
          for(int i = 0; i < j; i++) {
            const double h0 = H[i * kBM + j], h1 = H[(i + 1) * kBM + j];
            H[i * kBM + j] = cs[i] * h0 + sn[i] * h1;
            H[(i + 1) * kBM + j] = -sn[i] * h0 + cs[i] * h1;
          }
          const double p = H[j * kBM + j], qv = H[(j + 1) * kBM + j];
          double r = sqrt(p * p + qv * qv);
          if(r == 0.) r = 1e-300;
          cs[j] = p / r;
          sn[j] = qv / r;
          H[j * kBM + j] = r;
          H[(j + 1) * kBM + j] = 0.;
          g[j + 1] = -sn[j] * g[j];
          g[j] = cs[j] * g[j];
        }
        __syncthreads();
        rest = fabs(g[j + 1]);
        jend = j;
        if(rest <= tol || hn == 0.) {
          conv = true;
          break;
        }
        if(its >= maxits) break;
      }
      // y = H^{-1} g, x += M^{-1} (V y)
      if(tid == 0) {
        for(int i = jend; i >= 0; i--) {
          double s = g[i];
          for(int k = i + 1; k <= jend; k++) s -= H[i * kBM + k] * yv[k];
          yv[i] = s / H[i * kBM + i];
        }
      }
      __syncthreads();
      for(int r = tid; r < n; r += T) {
        double s = 0.;
        for(int i = 0; i <= jend; i++) s += yv[i] * V[(size_t)i * n + r];
        Z[r] = s;
      }
      __syncthreads();
      b_ilu<T, W>(a, inst, n, Z, Z, zlds, fifo);
      for(int r = tid; r < n; r += T) x[r] += Z[r];
      __syncthreads();
      if(conv) {
        st = 0;
        break;
      }
      if(its >= maxits) break;
      // restart: r = b - A x
      b_spmv(a, r0, n, x, V, T);
      __syncthreads();
      double rr = 0.;
      for(int r = tid; r < n; r += T) {
        const double v = b[r] - V[r];
        V[r] = v;
        rr += v * v;
      }
      beta = sqrt(block_sum_d<T>(rr, red));
      rest = beta;
      if(beta <= tol) {
        st = 0;
        break;
      }
    }
    if(tid == 0) {
      a.its[inst] = its;
      a.status[inst] = st;
      a.res0[inst] = beta0;
      a.res[inst] = rest;
      a.t1[inst] = wall_clock64();
    }
  }
}

__global__ void k_bgather(long nnz, const int *map, const double *a, double *b)
{
  long k = (long)blockIdx.x * 256 + threadIdx.x;
  if(k < nnz) b[k] = a[map[k]];
}
__global__ void k_bperm(int n, const int *rperm, const double *b, double *out)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n) out[i] = b[rperm[i]];
}
__global__ void k_bunperm(int n, const int *perm, const double *sol, double *x)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n) x[i] = sol[perm[i]];
}
// record values from the factor: val[i] = lu[k[i]] (0 for padding)
__global__ void k_pack_vals(long n, const int *k, const float *lu, float *val)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i < n) val[i] = k[i] >= 0 ? lu[k[i]] : 0.f;
}

__global__ void k_zero_ints(int n, int *p)
{
  int i = blockIdx.x * 256 + threadIdx.x;
  if(i < n) p[i] = 0;
}

}  // namespace
}  // namespace gdp

using namespace gdp;

struct gdp_batch {
  Params p;
  int ninst = 0, n = 0;
  std::vector<int> off;
  std::vector<int> cfg_threads, cfg_orth, cfg_order;
  bool prepared = false;
  const gdp_matrix *A = nullptr;
  uint64_t pattern_version = 0;
  // permuted system
  std::vector<int> perm, rperm, ib, jb, map;
  int *d_perm = nullptr, *d_rperm = nullptr, *d_map = nullptr, *d_off = nullptr;
  DevCSR Bm;
  IluFactor ilu;
  int *d_fl_base = nullptr, *d_fl_n = nullptr, *d_fl_ptr = nullptr;
  int *d_bl_base = nullptr, *d_bl_n = nullptr, *d_bl_ptr = nullptr;
  int W = 4;  // record width: most factor entries on one side of a row
  int *d_f_row = nullptr, *d_f_cnt = nullptr, *d_f_col = nullptr, *d_f_k = nullptr;
  int *d_b_row = nullptr, *d_b_cnt = nullptr, *d_b_col = nullptr, *d_b_k = nullptr, *d_b_dk = nullptr;
  int *d_ring = nullptr;  // per instance: sweeps through the LDS ring
  int4 *d_f_meta = nullptr, *d_b_meta = nullptr;
  float *d_f_val = nullptr, *d_b_val = nullptr, *d_b_dinv = nullptr;
  long *d_pool_off = nullptr;
  double *pool = nullptr, *rhs = nullptr, *sol = nullptr;
  // launch groups (threads, orthogonalisation)
  struct Group {
    int threads, orth;
    std::vector<int> inst;
    int *d_order = nullptr;
  };
  std::vector<Group> groups;
  int *d_queue = nullptr;
  int *d_its = nullptr, *d_status = nullptr;
  double *d_res0 = nullptr, *d_res = nullptr;
  unsigned long long *d_t0 = nullptr, *d_t1 = nullptr;
  std::vector<hipStream_t> side;
  hipEvent_t ev_fork = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
  std::vector<hipEvent_t> ev_join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  gdp_fe *fe = nullptr;
  gdp_vector *xv = nullptr, *bv = nullptr;
  gdp_matrix *Av = nullptr;
  double graph_ms = 0.;
  void release_graph()
  {
    if(exec) (void)hipGraphExecDestroy(exec);
    if(graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~gdp_batch()
  {
    release_graph();
    for(auto s : side) (void)hipStreamDestroy(s);
    for(auto e : ev_join) (void)hipEventDestroy(e);
    if(ev_fork) (void)hipEventDestroy(ev_fork);
    if(ev_t0) (void)hipEventDestroy(ev_t0);
    if(ev_t1) (void)hipEventDestroy(ev_t1);
    for(auto &g : groups) dfree(g.d_order);
    for(int **p : {&d_perm, &d_rperm, &d_map, &d_off, &d_fl_base, &d_fl_n, &d_fl_ptr, &d_bl_base, &d_bl_n,
                   &d_bl_ptr, &d_queue, &d_its, &d_status, &d_f_row, &d_f_cnt, &d_f_col, &d_f_k, &d_b_row,
                   &d_b_cnt, &d_b_col, &d_b_k, &d_b_dk, &d_ring})
      dfree(*p);
    dfree(d_f_meta);
    dfree(d_b_meta);
    dfree(d_f_val);
    dfree(d_b_val);
    dfree(d_b_dinv);
    dfree(d_pool_off);
    dfree(pool);
    dfree(rhs);
    dfree(sol);
    dfree(d_res0);
    dfree(d_res);
    dfree(d_t0);
    dfree(d_t1);
    Bm.release();
  }
};

namespace {

// rows of the sweep vector kept in LDS, per workgroup size (1, 2 or 3
// workgroups of 1024, 512, 256 threads per CU with 160 KB of LDS)
int zcap_rows(int threads) { return threads == 1024 ? 14336 : (threads == 512 ? 8192 : 4096); }
// entries of the LDS ring of larger instances: the largest power of two
// within the same budget
int ring_rows(int threads)
{
  int r = 1;
  while(2 * r <= zcap_rows(threads)) r *= 2;
  return r;
}

template <int T, int W> void launch_group(const BatchArgs &a, int blocks, hipStream_t s)
{
  const size_t lds = (size_t)a.zcap * sizeof(double);
  if(a.orth == 1) hipLaunchKernelGGL((k_batch_gmres<T, W, 1>), dim3(blocks), dim3(T), lds, s, a);
  else hipLaunchKernelGGL((k_batch_gmres<T, W, 0>), dim3(blocks), dim3(T), lds, s, a);
}

template <int T, int W> int allow_lds()
{
  GDP_HIP(hipFuncSetAttribute((const void *)k_batch_gmres<T, W, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              zcap_rows(T) * (int)sizeof(double)));
  GDP_HIP(hipFuncSetAttribute((const void *)k_batch_gmres<T, W, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              zcap_rows(T) * (int)sizeof(double)));
  return GDP_OK;
}

int batch_prepare(gdp_batch *B, const gdp_matrix *A)
{
  const int n = A->n;
  if(B->off.back() != n) return set_error(GDP_E_ARG, "batch covers %d rows, matrix has %d", B->off.back(), n);
  const int *ia = A->h_ia.data(), *ja = A->h_ja.data();
  // block-diagonal check and per-instance orderings
  std::vector<int> inst_of(n);
  for(int i = 0; i < B->ninst; i++)
    for(int r = B->off[i]; r < B->off[i + 1]; r++) inst_of[r] = i;
  for(int r = 0; r < n; r++) {
    if(ia[r + 1] <= ia[r]) return set_error(GDP_E_ZERO_ROW, "Zero row in matrix (row %d)", r + 1);
    for(int k = ia[r]; k < ia[r + 1]; k++)
      if(inst_of[ja[k]] != inst_of[r])
        return set_error(GDP_E_ARG, "entry (%d, %d) couples instances %d and %d", r, ja[k], inst_of[r],
                         inst_of[ja[k]]);
  }
  B->perm.assign(n, 0);
  B->rperm.assign(n, 0);
  for(int i = 0; i < B->ninst; i++) {
    const int r0 = B->off[i], ni = B->off[i + 1] - r0;
    std::vector<int> sia(ni + 1, 0), sja;
    for(int r = 0; r < ni; r++) {
      for(int k = ia[r0 + r]; k < ia[r0 + r + 1]; k++) sja.push_back(ja[k] - r0);
      sia[r + 1] = (int)sja.size();
    }
    std::vector<int> p(ni), rp(ni);
    for(int r = 0; r < ni; r++) p[r] = rp[r] = r;
    if(B->p.Renumbering_Technique == 1 && ni > 0) rcmk(ni, sia.data(), sja.data(), p, rp);
    if(B->cfg_order[i] == 1 && ni > 0) {
      std::vector<int> pib, pjb, pmap, col;
      permute_pattern(ni, sia.data(), sja.data(), p.data(), pib, pjb, pmap);
      const int ncol = greedy_colouring(ni, pib.data(), pjb.data(), col);
      std::vector<int> cnt(ncol + 1, 0), p2(ni);
      for(int r = 0; r < ni; r++) cnt[col[r] + 1]++;
      for(int k = 0; k < ncol; k++) cnt[k + 1] += cnt[k];
      for(int r = 0; r < ni; r++) p2[r] = cnt[col[r]]++;
      for(int r = 0; r < ni; r++) p[r] = p2[p[r]];
      for(int r = 0; r < ni; r++) rp[p[r]] = r;
    }
    for(int r = 0; r < ni; r++) {
      B->perm[r0 + r] = r0 + p[r];
      B->rperm[r0 + p[r]] = r0 + r;
    }
  }
  permute_pattern(n, ia, ja, B->perm.data(), B->ib, B->jb, B->map);
  const long nnz = (long)B->jb.size();
  GDP_TRY(dalloc(&B->d_perm, (size_t)n));
  GDP_TRY(dalloc(&B->d_rperm, (size_t)n));
  GDP_TRY(dalloc(&B->d_map, (size_t)std::max(1L, nnz)));
  GDP_TRY(dalloc(&B->d_off, (size_t)B->ninst + 1));
  GDP_TRY(h2d(B->d_perm, B->perm.data(), (size_t)n));
  GDP_TRY(h2d(B->d_rperm, B->rperm.data(), (size_t)n));
  GDP_TRY(h2d(B->d_map, B->map.data(), (size_t)nnz));
  GDP_TRY(h2d(B->d_off, B->off.data(), B->off.size()));
  B->Bm.n = n;
  B->Bm.nnz = nnz;
  GDP_TRY(dalloc(&B->Bm.ia, (size_t)n + 1));
  GDP_TRY(dalloc(&B->Bm.ja, (size_t)std::max(1L, nnz)));
  GDP_TRY(dalloc(&B->Bm.a, (size_t)std::max(1L, nnz)));
  GDP_TRY(h2d(B->Bm.ia, B->ib.data(), (size_t)n + 1));
  GDP_TRY(h2d(B->Bm.ja, B->jb.data(), (size_t)nnz));
  GDP_TRY(B->ilu.analyse(B->Bm, B->ib, B->jb));
  GDP_TRY(B->ilu.reserve(false));  // FP32 factor storage, as the reference
  // sweep levels per instance (levels never cross instances: blocks are independent)
  std::vector<int> flev, blev;
  lower_levels(n, B->ib.data(), B->jb.data(), flev);
  upper_levels(n, B->ib.data(), B->jb.data(), blev);
  // rows of instance i grouped by level: ptr[base[i] + l] .. ptr[base[i] + l + 1]
  // index `rows`; a sweep processes the levels in increasing order
  auto build = [&](const std::vector<int> &lev, std::vector<int> &base, std::vector<int> &cnt,
                   std::vector<int> &ptr, std::vector<int> &rows) {
    base.assign(B->ninst, 0);
    cnt.assign(B->ninst, 0);
    ptr.clear();
    rows.clear();
    for(int i = 0; i < B->ninst; i++) {
      const int r0 = B->off[i], r1 = B->off[i + 1];
      int nl = 0;
      for(int r = r0; r < r1; r++) nl = std::max(nl, lev[r] + 1);
      std::vector<std::vector<int>> by(nl);
      for(int r = r0; r < r1; r++) by[lev[r]].push_back(r);
      base[i] = (int)ptr.size();
      cnt[i] = nl;
      for(int l = 0; l < nl; l++) {
        ptr.push_back((int)rows.size());
        rows.insert(rows.end(), by[l].begin(), by[l].end());
      }
      ptr.push_back((int)rows.size());
    }
  };
  std::vector<int> fb, fn, fp, fr, bb, bn, bp, br;
  build(flev, fb, fn, fp, fr);
  build(blev, bb, bn, bp, br);
  // row records in slot order: local row, entry count, local columns and the
  // factor positions whose values the pack kernel copies after each factorisation
  std::vector<int> dpos(n, -1);
  int wmax = 1;
  for(int r = 0; r < n; r++) {
    for(int k = B->ib[r]; k < B->ib[r + 1]; k++)
      if(B->jb[k] == r) dpos[r] = k;
    if(dpos[r] < 0) return set_error(GDP_E_ZERO_PIVOT, "Zero pivot on line %d in ILU (missing diagonal)", r + 1);
    wmax = std::max({wmax, dpos[r] - B->ib[r], B->ib[r + 1] - dpos[r] - 1});
  }
  if(wmax > 8) return set_error(GDP_E_UNSUPPORTED, "batched sweeps: %d factor entries on one side of a row (max 8)", wmax);
  B->W = wmax <= 4 ? 4 : 8;
  const int W = B->W;
  // stream positions (per instance and sweep) and the LDS ring: an instance
  // too large for the LDS sweep vector reads its dependencies from a ring of
  // the last R results when no dependency lies R or more positions behind the
  // end of the level that needs it (R: a power of two within the LDS budget)
  std::vector<int> fpos(n, 0), bpos(n, 0), ringf(B->ninst, 0);
  auto positions = [&](const std::vector<int> &base, const std::vector<int> &cnt, const std::vector<int> &ptr,
                       const std::vector<int> &rows, std::vector<int> &pos) {
    for(int i = 0; i < B->ninst; i++)
      if(cnt[i] > 0)
        for(int q = ptr[base[i]]; q < ptr[base[i] + cnt[i]]; q++) pos[rows[q]] = q - ptr[base[i]];
  };
  positions(fb, fn, fp, fr, fpos);
  positions(bb, bn, bp, br, bpos);
  auto ring_ok = [&](int i, int R, const std::vector<int> &base, const std::vector<int> &cnt,
                     const std::vector<int> &ptr, const std::vector<int> &rows, bool lower,
                     const std::vector<int> &pos) {
    if(cnt[i] <= 0) return true;
    const int s0 = ptr[base[i]];
    for(int l = 0; l < cnt[i]; l++) {
      const int last = ptr[base[i] + l + 1] - s0 - 1;
      for(int q = ptr[base[i] + l]; q < ptr[base[i] + l + 1]; q++) {
        const int r = rows[q];
        const int k0 = lower ? B->ib[r] : dpos[r] + 1, k1 = lower ? dpos[r] : B->ib[r + 1];
        for(int k = k0; k < k1; k++)
          if(last - pos[B->jb[k]] >= R) return false;
      }
    }
    return true;
  };
  for(int i = 0; i < B->ninst; i++) {
    const int ni = B->off[i + 1] - B->off[i], zc = zcap_rows(B->cfg_threads[i]);
    if(ni <= zc) continue;
    const int R = ring_rows(B->cfg_threads[i]);
    ringf[i] = ring_ok(i, R, fb, fn, fp, fr, true, fpos) && ring_ok(i, R, bb, bn, bp, br, false, bpos);
  }
  auto records = [&](const std::vector<int> &rows, bool lower, std::vector<int> &rrow, std::vector<int> &rcnt,
                     std::vector<int> &rcol, std::vector<int> &rk, std::vector<int> *dk) {
    const size_t S = rows.size();
    rrow.assign(S, 0);
    rcnt.assign(S, 0);
    rcol.assign(S * W, 0);
    rk.assign(S * W, -1);
    if(dk) dk->assign(S, 0);
    for(size_t q = 0; q < S; q++) {
      const int r = rows[q], r0 = B->off[inst_of[r]];
      const int k0 = lower ? B->ib[r] : dpos[r] + 1, k1 = lower ? dpos[r] : B->ib[r + 1];
      rrow[q] = r - r0;
      rcnt[q] = k1 - k0;
      const bool rg = ringf[inst_of[r]] != 0;
      for(int k = k0; k < k1; k++) {
        rcol[q * W + (k - k0)] = rg ? (lower ? fpos : bpos)[B->jb[k]] : B->jb[k] - r0;
        rk[q * W + (k - k0)] = k;
      }
      if(dk) (*dk)[q] = dpos[r];
    }
  };
  std::vector<int> frow, fcnt, fcol, fk, brow, bcnt, bcol, bk, bdk;
  records(fr, true, frow, fcnt, fcol, fk, nullptr);
  records(br, false, brow, bcnt, bcol, bk, &bdk);
  // {row, count, level within the instance, 0} per slot
  auto metas = [&](const std::vector<int> &base, const std::vector<int> &cnt, const std::vector<int> &ptr,
                   const std::vector<int> &rrow, const std::vector<int> &rcnt, int4 **dst) {
    std::vector<int4> mt(std::max<size_t>(1, rrow.size()), make_int4(0, 0, 0, 0));
    for(int i = 0; i < B->ninst; i++)
      for(int l = 0; l < cnt[i]; l++)
        for(int q = ptr[base[i] + l]; q < ptr[base[i] + l + 1]; q++) mt[q] = make_int4(rrow[q], rcnt[q], l, 0);
    int rc = dalloc(dst, mt.size());
    if(!rc) rc = h2d(*dst, mt.data(), mt.size());
    return rc;
  };
  GDP_TRY(metas(fb, fn, fp, frow, fcnt, &B->d_f_meta));
  GDP_TRY(metas(bb, bn, bp, brow, bcnt, &B->d_b_meta));
  for(auto &pr : std::vector<std::pair<int **, std::vector<int> *>>{
        {&B->d_fl_base, &fb}, {&B->d_fl_n, &fn}, {&B->d_fl_ptr, &fp}, {&B->d_bl_base, &bb},
        {&B->d_bl_n, &bn}, {&B->d_bl_ptr, &bp}, {&B->d_f_row, &frow}, {&B->d_f_cnt, &fcnt},
        {&B->d_f_col, &fcol}, {&B->d_f_k, &fk}, {&B->d_b_row, &brow}, {&B->d_b_cnt, &bcnt},
        {&B->d_b_col, &bcol}, {&B->d_b_k, &bk}, {&B->d_b_dk, &bdk}, {&B->d_ring, &ringf}}) {
    GDP_TRY(dalloc(pr.first, std::max<size_t>(1, pr.second->size())));
    GDP_TRY(h2d(*pr.first, pr.second->data(), pr.second->size()));
  }
  GDP_TRY(dalloc(&B->d_f_val, std::max<size_t>(1, fk.size())));
  GDP_TRY(dalloc(&B->d_b_val, std::max<size_t>(1, bk.size())));
  GDP_TRY(dalloc(&B->d_b_dinv, std::max<size_t>(1, bdk.size())));
  // Krylov workspace: (m + 3) vectors per instance (basis, preconditioned
  // vector, the sweeps' staged inputs)
  const int m = std::max(2, std::min(B->p.Krylov_Size <= 1 ? 15 : B->p.Krylov_Size, kBM));
  B->p.Krylov_Size = m;
  std::vector<long> po(B->ninst);
  long tot = 0;
  for(int i = 0; i < B->ninst; i++) {
    po[i] = tot;
    tot += (long)(m + 3) * (B->off[i + 1] - B->off[i]);
  }
  GDP_TRY(dalloc(&B->d_pool_off, (size_t)B->ninst));
  GDP_TRY(h2d(B->d_pool_off, po.data(), po.size()));
  GDP_TRY(dalloc(&B->pool, (size_t)std::max(1L, tot)));
  GDP_TRY(dalloc(&B->rhs, (size_t)n));
  GDP_TRY(dalloc(&B->sol, (size_t)n));
  // launch groups, instances largest first
  std::map<std::pair<int, int>, int> gid;
  for(int i = 0; i < B->ninst; i++) {
    auto key = std::make_pair(B->cfg_threads[i], B->cfg_orth[i]);
    if(!gid.count(key)) {
      gid[key] = (int)B->groups.size();
      B->groups.push_back({key.first, key.second, {}, nullptr});
    }
    B->groups[gid[key]].inst.push_back(i);
  }
  for(auto &g : B->groups) {
    std::stable_sort(g.inst.begin(), g.inst.end(), [&](int x, int y) {
      return B->off[x + 1] - B->off[x] > B->off[y + 1] - B->off[y];
    });
    GDP_TRY(dalloc(&g.d_order, g.inst.size()));
    GDP_TRY(h2d(g.d_order, g.inst.data(), g.inst.size()));
  }
  GDP_TRY(dalloc(&B->d_queue, B->groups.size()));
  GDP_TRY(dalloc(&B->d_its, (size_t)B->ninst));
  GDP_TRY(dalloc(&B->d_status, (size_t)B->ninst));
  GDP_TRY(dalloc(&B->d_res0, (size_t)B->ninst));
  GDP_TRY(dalloc(&B->d_res, (size_t)B->ninst));
  GDP_TRY(dalloc(&B->d_t0, (size_t)B->ninst));
  GDP_TRY(dalloc(&B->d_t1, (size_t)B->ninst));
  for(size_t gi = 0; gi < B->groups.size(); gi++) {
    hipStream_t s;
    GDP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    B->side.push_back(s);
    hipEvent_t e;
    GDP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    B->ev_join.push_back(e);
  }
  GDP_HIP(hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming));
  GDP_HIP(hipEventCreate(&B->ev_t0));
  GDP_HIP(hipEventCreate(&B->ev_t1));
  GDP_TRY((allow_lds<256, 4>()));
  GDP_TRY((allow_lds<512, 4>()));
  GDP_TRY((allow_lds<1024, 4>()));
  GDP_TRY((allow_lds<256, 8>()));
  GDP_TRY((allow_lds<512, 8>()));
  GDP_TRY((allow_lds<1024, 8>()));
  B->A = A;
  B->pattern_version = A->pattern_version;
  B->prepared = true;
  return GDP_OK;
}



// Everything one run enqueues (captured once into the graph).
int batch_enqueue(gdp_batch *B, gdp_matrix *A, gdp_vector *b, gdp_vector *x)
{
  hipStream_t s = ctx().stream;
  const int n = B->n;
  const long nnz = B->Bm.nnz;
  if(B->fe) GDP_TRY(fe_generate_enqueue(B->fe, A, b));
  if(nnz)
    hipLaunchKernelGGL(k_bgather, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, nnz, B->d_map, A->csr.a,
                       B->Bm.a);
  GDP_TRY(B->ilu.factor_enqueue(B->Bm.a));
  {
    const long nf = (long)n * B->W;
    hipLaunchKernelGGL(k_pack_vals, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, nf, B->d_f_k,
                       B->ilu.lu32, B->d_f_val);
    hipLaunchKernelGGL(k_pack_vals, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, nf, B->d_b_k,
                       B->ilu.lu32, B->d_b_val);
    hipLaunchKernelGGL(k_pack_vals, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (long)n, B->d_b_dk,
                       B->ilu.lu32, B->d_b_dinv);
  }
  hipLaunchKernelGGL(k_bperm, dim3((n + 255) / 256), dim3(256), 0, s, n, B->d_rperm, b->d, B->rhs);
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(256), 0, s, (int)B->groups.size(), B->d_queue);
  GDP_HIP(hipEventRecord(B->ev_fork, s));
  int cu = ctx().cu_count;
  for(size_t gi = 0; gi < B->groups.size(); gi++) {
    auto &g = B->groups[gi];
    hipStream_t ss = B->side[gi];
    GDP_HIP(hipStreamWaitEvent(ss, B->ev_fork, 0));
    BatchArgs a{};
    a.m = B->p.Krylov_Size;
    a.maxits = B->p.Nb_Iter_Max;
    a.orth = g.orth;
    a.rtol = B->p.Stopping_Test;
    a.ia = B->Bm.ia;
    a.ja = B->Bm.ja;
    a.a = B->Bm.a;
    a.diag = B->ilu.diag;
    a.lu = B->ilu.lu32;
    a.off = B->d_off;
    a.fl_base = B->d_fl_base;
    a.fl_n = B->d_fl_n;
    a.fl_ptr = B->d_fl_ptr;

    a.bl_base = B->d_bl_base;
    a.bl_n = B->d_bl_n;
    a.bl_ptr = B->d_bl_ptr;
    a.f_row = B->d_f_row;
    a.f_cnt = B->d_f_cnt;
    a.f_col = B->d_f_col;
    a.f_val = B->d_f_val;
    a.b_row = B->d_b_row;
    a.b_cnt = B->d_b_cnt;
    a.b_col = B->d_b_col;
    a.b_val = B->d_b_val;
    a.b_dinv = B->d_b_dinv;
    a.f_meta = B->d_f_meta;
    a.b_meta = B->d_b_meta;
    a.zcap = zcap_rows(g.threads);
    a.ring = B->d_ring;
    a.rmask = ring_rows(g.threads) - 1;
    // GDP_BATCH_WALK=0: the level-by-level sweeps for every instance
    const char *ew = std::getenv("GDP_BATCH_WALK");
    a.walk = ew ? atoi(ew) != 0 : 1;
    a.pool_off = B->d_pool_off;
    a.pool = B->pool;
    a.rhs = B->rhs;
    a.sol = B->sol;
    a.order = g.d_order;
    a.count = (int)g.inst.size();
    a.queue = B->d_queue + gi;
    a.its = B->d_its;
    a.status = B->d_status;
    a.res0 = B->d_res0;
    a.res = B->d_res;
    a.t0 = B->d_t0;
    a.t1 = B->d_t1;
    // resident workgroups: 16 waves per CU at this kernel's register use,
    // shared between the groups (extra workgroups would find the queue empty)
    const int per_cu = std::max(1, 1024 / g.threads);
    const int blocks = std::max(1, std::min(a.count, cu * per_cu / (int)B->groups.size()));
    if(B->W == 4) {
      if(g.threads == 256) launch_group<256, 4>(a, blocks, ss);
      else if(g.threads == 512) launch_group<512, 4>(a, blocks, ss);
      else launch_group<1024, 4>(a, blocks, ss);
    }
    else {
      if(g.threads == 256) launch_group<256, 8>(a, blocks, ss);
      else if(g.threads == 512) launch_group<512, 8>(a, blocks, ss);
      else launch_group<1024, 8>(a, blocks, ss);
    }
    GDP_HIP(hipGetLastError());
    GDP_HIP(hipEventRecord(B->ev_join[gi], ss));
    GDP_HIP(hipStreamWaitEvent(s, B->ev_join[gi], 0));
  }
  hipLaunchKernelGGL(k_bunperm, dim3((n + 255) / 256), dim3(256), 0, s, n, B->d_perm, B->sol, x->d);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

}  // namespace

extern "C" int gdp_batch_create(const gdp_solver *params, int ninst, const int *row_offsets,
                                const int *config, gdp_batch **out)
{
  if(!params || !row_offsets || !out || ninst < 1) return set_error(GDP_E_ARG, "gdp_batch_create");
  GDP_TRY(ensure_ready());
  for(int i = 0; i < ninst; i++)
    if(row_offsets[i + 1] < row_offsets[i] || row_offsets[0] != 0)
      return set_error(GDP_E_ARG, "row offsets must start at 0 and not decrease");
  auto *B = new gdp_batch;
  double v;
  const char *keys[] = {"Algorithm", "Preconditioner", "Krylov_Size", "Nb_Iter_Max", "Stopping_Test",
                        "Renumbering_Technique", "Scaling"};
  for(const char *k : keys) {
    if(gdp_solver_get(params, k, &v) < 0) {
      delete B;
      return GDP_E_ARG;
    }
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    params_set(B->p, k, buf);
  }
  if(B->p.Algorithm != 8 || B->p.Preconditioner != 6 || B->p.Scaling != 0) {
    delete B;
    return set_error(GDP_E_UNSUPPORTED, "batched path: GMRES (8) with ILU(0) (6), no scaling");
  }
  if(B->p.Krylov_Size > kBM) {
    delete B;
    return set_error(GDP_E_UNSUPPORTED, "batched path: Krylov_Size at most %d", kBM);
  }
  B->ninst = ninst;
  B->off.assign(row_offsets, row_offsets + ninst + 1);
  B->n = B->off.back();
  B->cfg_threads.assign(ninst, 256);
  B->cfg_orth.assign(ninst, 1);
  B->cfg_order.assign(ninst, 0);
  if(config)
    for(int i = 0; i < ninst; i++) {
      const int t = config[3 * i], o = config[3 * i + 1], ord = config[3 * i + 2];
      if((t != 256 && t != 512 && t != 1024) || (o != 0 && o != 1) || (ord != 0 && ord != 1)) {
        delete B;
        return set_error(GDP_E_ARG, "instance %d: config (%d, %d, %d)", i, t, o, ord);
      }
      B->cfg_threads[i] = t;
      B->cfg_orth[i] = o;
      B->cfg_order[i] = ord;
    }
  *out = B;
  return GDP_OK;
}

extern "C" int gdp_batch_destroy(gdp_batch *B)
{
  if(!B) return GDP_OK;
  if(ctx().ready) (void)sync();
  delete B;
  return GDP_OK;
}

// Assemble (when fe is given) and solve every instance: the first call
// analyses the pattern and records the graph, later calls replay it.
extern "C" int gdp_batch_run(gdp_batch *B, gdp_fe *fe, gdp_matrix *A, gdp_vector *b, gdp_vector *x)
{
  if(!B || !A || !b || !x) return set_error(GDP_E_ARG, "gdp_batch_run: null argument");
  if(A->n != B->n || b->n != B->n || x->n != B->n) return set_error(GDP_E_ARG, "gdp_batch_run: sizes");
  hipStream_t s = ctx().stream;
  if(fe && !B->prepared) GDP_TRY(gdp_fe_generate(fe, A, b));  // uploads and adopts the pattern
  if(A->csr.n != B->n) return set_error(GDP_E_STATE, "gdp_batch_run: matrix not assembled");
  if(!B->prepared || B->A != A || B->pattern_version != A->pattern_version) {
    if(B->prepared) return set_error(GDP_E_STATE, "gdp_batch_run: the matrix pattern changed");
    GDP_TRY(batch_prepare(B, A));
  }
  if(!B->exec || B->fe != fe || B->Av != A || B->bv != b || B->xv != x) {
    B->release_graph();
    B->fe = fe;
    B->Av = A;
    B->bv = b;
    B->xv = x;
    GDP_TRY(sync());
    GDP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(B, A, b, x);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if(rc < 0) {
      if(g) (void)hipGraphDestroy(g);
      return rc;
    }
    GDP_HIP(e);
    B->graph = g;
    GDP_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  }
  GDP_HIP(hipEventRecord(B->ev_t0, s));
  GDP_HIP(hipGraphLaunch(B->exec, s));
  GDP_HIP(hipEventRecord(B->ev_t1, s));
  GDP_TRY(sync());
  float ms = 0.f;
  GDP_HIP(hipEventElapsedTime(&ms, B->ev_t0, B->ev_t1));
  B->graph_ms = ms;
  int row = 0;
  GDP_TRY(B->ilu.factor_status(&row));
  if(row) return set_error(GDP_E_ZERO_PIVOT, "Zero pivot on line %d in ILU", row);
  return GDP_OK;
}

// Per-instance results of the last run: iterations, status (0 converged,
// -1 iteration limit), |r0|, the estimated final residual norm, and the
// instance's start and end on the device clock (microseconds from the
// earliest start); graph_ms: device time of the whole graph.
extern "C" int gdp_batch_info(const gdp_batch *B, int *its, int *status, double *res0, double *res,
                              double *t_begin_us, double *t_end_us, double *graph_ms)
{
  if(!B || !B->prepared) return set_error(GDP_E_STATE, "gdp_batch_info before a run");
  const size_t n = (size_t)B->ninst;
  if(its) GDP_TRY(d2h(its, B->d_its, n));
  if(status) GDP_TRY(d2h(status, B->d_status, n));
  if(res0) GDP_TRY(d2h(res0, B->d_res0, n));
  if(res) GDP_TRY(d2h(res, B->d_res, n));
  if(t_begin_us || t_end_us) {
    std::vector<unsigned long long> a(n), b(n);
    GDP_TRY(d2h(a.data(), B->d_t0, n));
    GDP_TRY(d2h(b.data(), B->d_t1, n));
    int khz = 100000;
    (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx().device);
    if(khz <= 0) khz = 100000;
    unsigned long long base = *std::min_element(a.begin(), a.end());
    for(size_t i = 0; i < n; i++) {
      if(t_begin_us) t_begin_us[i] = (double)(a[i] - base) * 1e3 / khz;
      if(t_end_us) t_end_us[i] = (double)(b[i] - base) * 1e3 / khz;
    }
  }
  if(graph_ms) *graph_ms = B->graph_ms;
  return GDP_OK;
}
