// Learned preconditioner: a 4-level convolutional U-Net on the nodal grid,
// run and trained with BF16 MFMA kernels (FP32 accumulation) on gfx950.
//
// This component has no counterpart in the reference; it is used as the
// (nonlinear) right preconditioner of flexible GMRES (iters.f:2522-2807
// restated in krylov.hip), with ILU(0)-GMRES as the fallback.
//
// Network (channels 16-32-64-128, 3x3 convolutions, ReLU):
//   E0: conv(in 2 -> 16), conv(16 -> 16)                          -> s0  (H)
//   E1: conv(pool s0 -> 32), conv(32 -> 32)                        -> s1  (H/2)
//   E2: conv(pool s1 -> 64), conv(64 -> 64)                        -> s2  (H/4)
//   E3: conv(pool s2 -> 128), conv(128 -> 128)                     -> s3  (H/8)
//   D2: conv(up s3 ++ s2 -> 64), conv(64 -> 64)                    -> d2
//   D1: conv(up d2 ++ s1 -> 32), conv(32 -> 32)                    -> d1
//   D0: conv(up d1 ++ s0 -> 16), conv(16 -> 16)                    -> d0
//   out: 1x1 conv(16 -> 1), no activation                          -> z
// pool = 2x2 average, up = nearest neighbour, ++ = channel concatenation;
// both are fused into the operand loaders of the next convolution.
//
// The same network runs on 3D grids (D x H x W voxels): 3x3x3 convolutions
// (27 taps), 2x2x2 average pooling and nearest upsampling in all three
// directions.  A 2D grid is the case D = 1 with 9 taps.
//
// Layout: activations NHWC in BF16 with channels padded to a multiple of 8,
// so that one lane loads 8 consecutive channels (16 bytes) of one pixel.
// A 3x3 convolution is an implicit GEMM: rows = pixels, columns = output
// channels, reduction index k = tap*Cin_pad + ci (padded to a multiple of
// 32).  MFMA 16x16x32 BF16 tiles: operand lane l holds row (l % 16) and
// reduction entries 8*(l/16)..+7; the accumulator lane l holds column l % 16,
// rows 4*(l/16)..+3.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>

#include "krylov.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

namespace gdp {

namespace {

enum LoadMode { M_PLAIN = 0, M_POOL = 1, M_UPCAT = 2 };

__device__ inline float bf2f(u16 b)
{
  unsigned int u = ((unsigned int)b) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__device__ __host__ inline u16 f2bf(float f)
{
  unsigned int u;
  __builtin_memcpy(&u, &f, 4);
  if((u & 0x7f800000u) == 0x7f800000u) return (u16)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);                            // round to nearest even
  return (u16)(u >> 16);
}

struct ConvArgs {
  int B, H, W;        // output (= logical input) spatial size
  int D, d3;          // depth (1 on 2D grids); d3: 3D grid (27 taps, 2x2x2 pooling)
  int cin_pad, cout;  // logical input channels (padded), output channels
  int K_pad;          // reduction length, multiple of 32
  int c1;             // UPCAT: channels of the coarse tensor
  const u16 *in, *in2;
  const u16 *wt;      // [cout][K_pad] (forward) or [cin][K2_pad] (adjoint)
  const float *bias;  // may be null
  u16 *out_bf;        // BF16 output with bias + ReLU, or
  float *out_f;       // FP32 output (plain)
  int relu;
  const int *gate;
  int gmode;
  int kchunk;         // split-K: reduction entries per blockIdx.z (multiple of 32)
  float *split;       // split-K partials [gridDim.z][P][cout] (null: fused epilogue)
};

__device__ inline bool unet_gated(const int *gate, int gmode)
{
  if(!gmode || !gate) return false;
  if(gmode == 1) return gate[0] != 0;
  return (gate[0] | gate[3]) != 0;
}

// one logical input element at voxel (b, z, y, x) (zeros outside the grid)
template <int MODE>
__device__ inline float load1(const ConvArgs &a, int b, int z, int y, int x, int ci)
{
  if(z < 0 || y < 0 || x < 0 || z >= a.D || y >= a.H || x >= a.W) return 0.f;
  if(MODE == M_PLAIN) return bf2f(a.in[((((long)b * a.D + z) * a.H + y) * a.W + x) * a.cin_pad + ci]);
  if(MODE == M_POOL) {
    const long W2 = 2 * a.W, H2 = 2 * a.H, D2 = a.d3 ? 2 * a.D : 1;
    const long cin = a.cin_pad;
    const u16 *p = a.in + ((((long)b * D2 + (a.d3 ? 2 * z : 0)) * H2 + 2 * y) * W2 + 2 * x) * cin + ci;
    float s = bf2f(p[0]) + bf2f(p[cin]) + bf2f(p[W2 * cin]) + bf2f(p[W2 * cin + cin]);
    if(a.d3) {
      const u16 *p1 = p + H2 * W2 * cin;
      s += bf2f(p1[0]) + bf2f(p1[cin]) + bf2f(p1[W2 * cin]) + bf2f(p1[W2 * cin + cin]);
      return bf2f(f2bf(0.125f * s));
    }
    return bf2f(f2bf(0.25f * s));
  }
  if(ci < a.c1) {
    const int Wc = a.W >> 1, Hc = a.H >> 1, Dc = a.d3 ? a.D >> 1 : 1;
    const int zc = a.d3 ? z >> 1 : 0;
    return bf2f(a.in[((((long)b * Dc + zc) * Hc + (y >> 1)) * Wc + (x >> 1)) * a.c1 + ci]);
  }
  const int c2 = a.cin_pad - a.c1;
  return bf2f(a.in2[((((long)b * a.D + z) * a.H + y) * a.W + x) * c2 + (ci - a.c1)]);
}

__device__ inline bf16x8 as_bf8(uint4 v)
{
  bf16x8 r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

// Implicit-GEMM convolution: a wave computes 16*TM pixels x 16*TN channels
// over the reduction range of its blockIdx.z; a block of 4 waves takes 4
// consecutive pixel tiles.  Without split-K the epilogue adds the bias,
// applies ReLU and stores; with split-K it stores FP32 partials.
//
// The reduction position of a lane's 8-channel group (tap, channel) and the
// pixel and weight addresses are carried from one 32-wide k-step to the next
// by additions only: no division or address rebuild inside the loop.
// tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1) (dz = 0 and tap < 9 on 2D grids)
__device__ inline int tap_dz(int tap) { return ((tap * 57) >> 9) - 1; }  // tap / 9 - 1, tap < 27
__device__ inline int tap_t9(int tap) { return tap - 9 * ((tap * 57) >> 9); }
__device__ inline int tap_dy(int tap) { return ((tap_t9(tap) * 43) >> 7) - 1; }  // (tap % 9) / 3 - 1
__device__ inline int tap_dx(int tap) { return tap_t9(tap) - 3 * ((tap_t9(tap) * 43) >> 7) - 1; }

// 8 input channels (ci0..ci0+7, logical input of the convolution) of the
// voxel at offset (dz, dy, dx) from the lane's output voxel; zeros outside.
// pix: PLAIN/UPCAT voxel index ((b*D + z)*H + y)*W + x, POOL fine voxel of
// (2z, 2y, 2x) (z = 0 on 2D grids); cb: b * (D/2) (UPCAT coarse planes; b on
// 2D grids).
template <int MODE>
__device__ inline uint4 conv_load_a(const ConvArgs &a, long pix, int pz, int py, int px, int cb, int dz, int dy,
                                    int dx, int ci0, bool ok)
{
  uint4 v = {0u, 0u, 0u, 0u};
  const int H = a.H, W = a.W, cin = a.cin_pad;
  const int z = pz + dz, y = py + dy, x = px + dx;
  if(!(ok && (unsigned)z < (unsigned)a.D && (unsigned)y < (unsigned)H && (unsigned)x < (unsigned)W)) return v;
  if(MODE == M_PLAIN) return *(const uint4 *)(a.in + (pix + ((long)dz * H + dy) * W + dx) * cin + ci0);
  if(MODE == M_POOL) {
    const long W2 = 2 * W, H2 = 2 * H;
    const u16 *p0 = a.in + (pix + (2L * dz * H2 + 2L * dy) * W2 + 2 * dx) * cin + ci0;
    uint4 q[8];
    q[0] = *(const uint4 *)p0;
    q[1] = *(const uint4 *)(p0 + cin);
    q[2] = *(const uint4 *)(p0 + W2 * cin);
    q[3] = *(const uint4 *)(p0 + W2 * cin + cin);
    if(a.d3) {
      const u16 *p1 = p0 + H2 * W2 * cin;
      q[4] = *(const uint4 *)p1;
      q[5] = *(const uint4 *)(p1 + cin);
      q[6] = *(const uint4 *)(p1 + W2 * cin);
      q[7] = *(const uint4 *)(p1 + W2 * cin + cin);
    }
    u16 *o = (u16 *)&v;
#pragma unroll
    for(int c = 0; c < 8; c++) {
      float s = bf2f(((const u16 *)&q[0])[c]) + bf2f(((const u16 *)&q[1])[c]) + bf2f(((const u16 *)&q[2])[c]) +
                bf2f(((const u16 *)&q[3])[c]);
      if(a.d3)
        s += bf2f(((const u16 *)&q[4])[c]) + bf2f(((const u16 *)&q[5])[c]) + bf2f(((const u16 *)&q[6])[c]) +
             bf2f(((const u16 *)&q[7])[c]);
      o[c] = f2bf((a.d3 ? 0.125f : 0.25f) * s);
    }
    return v;
  }
  // upsampled coarse channels or skip channels: one load from a selected address
  const bool coarse = ci0 < a.c1;
  const int zc = a.d3 ? z >> 1 : 0;
  const u16 *pc = a.in + (((long)(cb + zc) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * a.c1 + ci0;
  const u16 *pf = a.in2 + (pix + ((long)dz * H + dy) * W + dx) * (cin - a.c1) + (ci0 - a.c1);
  return *(const uint4 *)(coarse ? pc : pf);
}

template <int MODE, int TM, int TN>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a)
{
  if(unet_gated(a.gate, a.gmode)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long P = (long)a.B * a.D * a.H * a.W;
  const long pbase = ((long)blockIdx.x * 4 + wave) * (TM * 16);
  const int cobase = blockIdx.y * 16 * TN;
  const int r16 = lane & 15, kq = lane >> 4;
  const int H = a.H, W = a.W, cin = a.cin_pad;
  const int taps = a.d3 ? 27 : 9, t0 = a.d3 ? 0 : 9;  // t0: 2D taps are the dz = 0 plane
  int pz[TM], py[TM], px[TM];
  long pix[TM];   // PLAIN/UPCAT: voxel index; POOL: fine voxel of (2z, 2y, 2x)
  int cb[TM];     // UPCAT: b * (D/2) (coarse plane base)
  bool pok[TM];
#pragma unroll
  for(int t = 0; t < TM; t++) {
    long p = pbase + t * 16 + r16;
    pok[t] = p < P;
    long q = pok[t] ? p : 0;
    px[t] = (int)(q % W);
    q /= W;
    py[t] = (int)(q % H);
    q /= H;
    pz[t] = (int)(q % a.D);
    const int b = (int)(q / a.D);
    if(MODE == M_POOL)
      pix[t] = a.d3 ? (((long)b * 2 * a.D + 2 * pz[t]) * (2 * H) + 2 * py[t]) * (2 * W) + 2 * px[t]
                    : ((long)b * 2 * H + 2 * py[t]) * (2 * W) + 2 * px[t];
    else pix[t] = pok[t] ? p : 0;
    cb[t] = a.d3 ? b * (a.D >> 1) : b;
  }
  f32x4 acc[TM][TN];
#pragma unroll
  for(int t = 0; t < TM; t++)
#pragma unroll
    for(int n = 0; n < TN; n++) acc[t][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  const int kbeg = blockIdx.z * a.kchunk;
  const int kend = min(a.K_pad, kbeg + a.kchunk);
  int tap = (kbeg + 8 * kq) / cin;
  int ci0 = kbeg + 8 * kq - tap * cin;
  const u16 *wp[TN];
  bool wok[TN];
#pragma unroll
  for(int n = 0; n < TN; n++) {
    const int co = cobase + n * 16 + r16;
    wok[n] = co < a.cout;
    wp[n] = a.wt + (long)(wok[n] ? co : 0) * a.K_pad + kbeg + 8 * kq;
  }
  // PF k-steps per round, all loads of a round issued before its MFMAs
  constexpr int PF = (TM * TN <= 4) ? 4 : 2;
  for(int kk = kbeg; kk < kend; kk += 32 * PF) {
    uint4 av[PF][TM], bv[PF][TN];
#pragma unroll
    for(int j = 0; j < PF; j++) {
      const bool on = kk + 32 * j < kend;
      const bool kvalid = on && tap < taps;
      const int dz = tap_dz(tap + t0), dy = tap_dy(tap + t0), dx = tap_dx(tap + t0);
#pragma unroll
      for(int t = 0; t < TM; t++)
        av[j][t] = conv_load_a<MODE>(a, pix[t], pz[t], py[t], px[t], cb[t], dz, dy, dx, ci0, kvalid && pok[t]);
#pragma unroll
      for(int n = 0; n < TN; n++) {
        bv[j][n] = (on && wok[n]) ? *(const uint4 *)wp[n] : zero4;
        wp[n] += on ? 32 : 0;
      }
      if(on) {
        ci0 += 32;
        while(ci0 >= cin) {
          ci0 -= cin;
          tap++;
        }
      }
    }
#pragma unroll
    for(int j = 0; j < PF; j++)
#pragma unroll
      for(int t = 0; t < TM; t++)
#pragma unroll
        for(int n = 0; n < TN; n++)
          acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[j][t]), as_bf8(bv[j][n]), acc[t][n], 0, 0, 0);
  }
  // epilogue: lane holds column (l % 16), rows 4*(l/16)..+3
#pragma unroll
  for(int t = 0; t < TM; t++)
#pragma unroll
    for(int n = 0; n < TN; n++) {
      const int co = cobase + n * 16 + r16;
      if(co >= a.cout) continue;
      if(a.split) {
        float *dst = a.split + (long)blockIdx.z * P * a.cout;
#pragma unroll
        for(int i = 0; i < 4; i++) {
          const long p = pbase + t * 16 + 4 * kq + i;
          if(p < P) dst[p * a.cout + co] = acc[t][n][i];
        }
        continue;
      }
      const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
      for(int i = 0; i < 4; i++) {
        const long p = pbase + t * 16 + 4 * kq + i;
        if(p >= P) continue;
        float v = acc[t][n][i] + bias;
        if(a.relu) v = v > 0.f ? v : 0.f;
        if(a.out_bf) a.out_bf[p * a.cout + co] = f2bf(v);
        else a.out_f[p * a.cout + co] = v;
      }
    }
}

// LDS-staged variant for grids whose tiles divide them: a workgroup computes
// an output tile of 16 x 16 pixels (2D) or 4 x 4 x 16 voxels (3D) times
// 16*TN channels.  The tile's input halo (the convolution's logical input:
// pooled, or upsampled and concatenated, as the loader forms it -- once per
// voxel instead of once per tap) is staged in LDS, CC channels at a time;
// every tap then reads its A fragments from LDS.  Wave w owns the tile's
// pixel rows 4w .. 4w + 3 (rows along x).  The reduction runs chunk by chunk,
// tap-major within a chunk.
constexpr int kTileX = 16;
__host__ __device__ inline int tile_ty(int d3) { return d3 ? 4 : 16; }
__host__ __device__ inline int tile_tz(int d3) { return d3 ? 4 : 1; }

template <int MODE, int TN>
__global__ void __launch_bounds__(256) k_conv_lds(ConvArgs a, int CC)
{
  extern __shared__ uint4 hal[];
  if(unet_gated(a.gate, a.gmode)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kq = lane >> 4;
  const int d3 = a.d3, TY = tile_ty(d3), TZ = tile_tz(d3);
  const int HX = kTileX + 2, HY = TY + 2, HZ = d3 ? TZ + 2 : 1;
  const int nx = a.W / kTileX, ny = a.H / TY, nz = d3 ? a.D / TZ : 1;
  long t = blockIdx.x;
  const int tx = (int)(t % nx);
  t /= nx;
  const int ty = (int)(t % ny);
  t /= ny;
  const int tz = (int)(t % nz);
  const int b = (int)(t / nz);
  const int x0 = tx * kTileX, y0 = ty * TY, z0 = tz * TZ;
  const int cobase = blockIdx.y * 16 * TN;
  const int taps = d3 ? 27 : 9, t0 = d3 ? 0 : 9;
  const int cin = a.cin_pad;
  const int VS = CC / 8 + 1;  // voxel stride in LDS (16-byte units, one of padding)
  const int nvox = HZ * HY * HX;
  f32x4 acc[4][TN];
#pragma unroll
  for(int i = 0; i < 4; i++)
#pragma unroll
    for(int n = 0; n < TN; n++) acc[i][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const u16 *wp[TN];
  bool wok[TN];
#pragma unroll
  for(int n = 0; n < TN; n++) {
    const int co = cobase + n * 16 + r16;
    wok[n] = co < a.cout;
    wp[n] = a.wt + (long)(wok[n] ? co : 0) * a.K_pad;
  }
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  const int ksteps = (taps * CC + 31) / 32;
  for(int c0 = 0; c0 < cin; c0 += CC) {
    __syncthreads();
    // stage the halo: voxel v (x fastest), CC / 8 channel octets each
    for(int i = threadIdx.x; i < nvox * (CC / 8); i += 256) {
      const int v = i / (CC / 8), g = i - v * (CC / 8);
      const int hx = v % HX, hy = (v / HX) % HY, hz = v / (HX * HY);
      const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = d3 ? z0 - 1 + hz : 0;
      const bool in = (unsigned)x < (unsigned)a.W && (unsigned)y < (unsigned)a.H && (unsigned)z < (unsigned)a.D;
      uint4 o = zero4;
      if(in) {
        const long pix = MODE == M_POOL ? (d3 ? (((long)b * 2 * a.D + 2 * z) * (2 * a.H) + 2 * y) * (2 * a.W) + 2 * x
                                              : ((long)b * 2 * a.H + 2 * y) * (2 * a.W) + 2 * x)
                                        : (((long)b * a.D + z) * a.H + y) * a.W + x;
        o = conv_load_a<MODE>(a, pix, z, y, x, d3 ? b * (a.D >> 1) : b, 0, 0, 0, c0 + 8 * g, true);
      }
      hal[v * VS + g] = o;
    }
    __syncthreads();
    for(int s = 0; s < ksteps; s++) {
      const int k = 32 * s + 8 * kq;  // this lane's 8 reduction entries
      const int tap = k / CC, ci = k - tap * CC;
      const bool kvalid = tap < taps;
      const int tt = tap + t0;
      const int dz = tap_dz(tt), dy = tap_dy(tt), dx = tap_dx(tt);
      uint4 bv[TN];
#pragma unroll
      for(int n = 0; n < TN; n++) bv[n] = (kvalid && wok[n]) ? *(const uint4 *)(wp[n] + tap * cin + c0 + ci) : zero4;
#pragma unroll
      for(int tr = 0; tr < 4; tr++) {
        const int rr = 4 * wave + tr;  // tile row: (z, y) = (rr / TY, rr % TY) in 3D, (0, rr) in 2D
        const int zz = d3 ? rr / TY : 0, yy = d3 ? rr % TY : rr;
        const int v = ((zz + (d3 ? 1 + dz : 0)) * HY + (yy + 1 + dy)) * HX + (r16 + 1 + dx);
        const uint4 av = kvalid ? hal[v * VS + ci / 8] : zero4;
#pragma unroll
        for(int n = 0; n < TN; n++)
          acc[tr][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av), as_bf8(bv[n]), acc[tr][n], 0, 0, 0);
      }
    }
  }
  // epilogue: lane holds channel r16 of pixels x0 + 4 kq + i of row tr
#pragma unroll
  for(int tr = 0; tr < 4; tr++) {
    const int rr = 4 * wave + tr;
    const int zz = d3 ? rr / TY : 0, yy = d3 ? rr % TY : rr;
    const long prow = (((long)b * a.D + z0 + zz) * a.H + y0 + yy) * a.W + x0;
#pragma unroll
    for(int n = 0; n < TN; n++) {
      const int co = cobase + n * 16 + r16;
      if(co >= a.cout) continue;
      const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
      for(int i = 0; i < 4; i++) {
        const long p = prow + 4 * kq + i;
        float v = acc[tr][n][i] + bias;
        if(a.relu) v = v > 0.f ? v : 0.f;
        if(a.out_bf) a.out_bf[p * a.cout + co] = f2bf(v);
        else a.out_f[p * a.cout + co] = v;
      }
    }
  }
}

// Small-image variant: the 4 waves of a block share one tile of 16 pixels x
// 16*TN channels and split the reduction into 4 contiguous ranges; their
// accumulators are added in LDS in wave order (deterministic) and wave 0
// stores.  Used when one wave per tile would leave most of the GPU idle.
template <int MODE, int TN>
__global__ void __launch_bounds__(256) k_conv_ks(ConvArgs a)
{
  if(unet_gated(a.gate, a.gmode)) return;
  __shared__ f32x4 red[3][TN][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long P = (long)a.B * a.D * a.H * a.W;
  const long pbase = (long)blockIdx.x * 16;
  const int cobase = blockIdx.y * 16 * TN;
  const int r16 = lane & 15, kq = lane >> 4;
  const int H = a.H, W = a.W, cin = a.cin_pad;
  const int taps = a.d3 ? 27 : 9, t0 = a.d3 ? 0 : 9;
  const long p = pbase + r16;
  const bool pok = p < P;
  long q = pok ? p : 0;
  const int px = (int)(q % W);
  q /= W;
  const int py = (int)(q % H);
  q /= H;
  const int pz = (int)(q % a.D);
  const int b = (int)(q / a.D);
  const long pix = MODE == M_POOL ? (a.d3 ? (((long)b * 2 * a.D + 2 * pz) * (2 * H) + 2 * py) * (2 * W) + 2 * px
                                          : ((long)b * 2 * H + 2 * py) * (2 * W) + 2 * px)
                                  : (pok ? p : 0);
  const int cb = a.d3 ? b * (a.D >> 1) : b;
  const int nsteps = a.K_pad / 32;
  const int s0 = wave * nsteps / 4, s1 = (wave + 1) * nsteps / 4;
  f32x4 acc[TN];
#pragma unroll
  for(int n = 0; n < TN; n++) acc[n] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  int tap = (32 * s0 + 8 * kq) / cin;
  int ci0 = 32 * s0 + 8 * kq - tap * cin;
  const u16 *wp[TN];
  bool wok[TN];
#pragma unroll
  for(int n = 0; n < TN; n++) {
    const int co = cobase + n * 16 + r16;
    wok[n] = co < a.cout;
    wp[n] = a.wt + (long)(wok[n] ? co : 0) * a.K_pad + 32 * s0 + 8 * kq;
  }
  // up to PF k-steps per round: every load of the round is issued before the
  // first MFMA, so a round costs one memory latency
  constexpr int PF = 8;
  for(int c0 = s0; c0 < s1; c0 += PF) {
    uint4 av[PF], bv[PF][TN];
#pragma unroll
    for(int j = 0; j < PF; j++) {
      const bool on = c0 + j < s1;
      av[j] = conv_load_a<MODE>(a, pix, pz, py, px, cb, tap_dz(tap + t0), tap_dy(tap + t0), tap_dx(tap + t0), ci0,
                                on && tap < taps && pok);
#pragma unroll
      for(int n = 0; n < TN; n++) {
        bv[j][n] = (on && wok[n]) ? *(const uint4 *)wp[n] : zero4;
        wp[n] += on ? 32 : 0;
      }
      if(on) {
        ci0 += 32;
        while(ci0 >= cin) {
          ci0 -= cin;
          tap++;
        }
      }
    }
#pragma unroll
    for(int j = 0; j < PF; j++)
#pragma unroll
      for(int n = 0; n < TN; n++)
        acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[j]), as_bf8(bv[j][n]), acc[n], 0, 0, 0);
  }
  if(wave > 0)
#pragma unroll
    for(int n = 0; n < TN; n++) red[wave - 1][n][lane] = acc[n];
  __syncthreads();
  if(wave > 0) return;
#pragma unroll
  for(int w = 0; w < 3; w++)
#pragma unroll
    for(int n = 0; n < TN; n++) acc[n] += red[w][n][lane];
#pragma unroll
  for(int n = 0; n < TN; n++) {
    const int co = cobase + n * 16 + r16;
    if(co >= a.cout) continue;
    const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for(int i = 0; i < 4; i++) {
      const long pp = pbase + 4 * kq + i;
      if(pp >= P) continue;
      float v = acc[n][i] + bias;
      if(a.relu) v = v > 0.f ? v : 0.f;
      if(a.out_bf) a.out_bf[pp * a.cout + co] = f2bf(v);
      else a.out_f[pp * a.cout + co] = v;
    }
  }
}

// split-K epilogue: sum the partials in a fixed order, bias, ReLU, store
__global__ void __launch_bounds__(256) k_conv_split_sum(ConvArgs a, int nsplit)
{
  if(unet_gated(a.gate, a.gmode)) return;
  const long P = (long)a.B * a.D * a.H * a.W;
  const long n = P * a.cout;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i >= n) return;
  float v = 0.f;
  for(int z = 0; z < nsplit; z++) v += a.split[(long)z * n + i];
  if(a.bias) v += a.bias[i % a.cout];
  if(a.relu) v = v > 0.f ? v : 0.f;
  if(a.out_bf) a.out_bf[i] = f2bf(v);
  else a.out_f[i] = v;
}

// Weight gradient: dW[co][k] = sum_p G[p][co] * Xcol[p][k] (k in the forward
// weight layout), split over pixel chunks; rows = co, columns = k, the MFMA
// reduction runs over pixels.  Partial results per chunk.
struct WgradArgs {
  ConvArgs x;          // logical input (MODE loader) and sizes
  const u16 *g;        // [P][cout] BF16 output gradient (ReLU-masked)
  float *part;         // [nchunks][cout][K_pad]
  int chunk;           // pixels per chunk (multiple of 32)
};

template <int MODE>
__global__ void __launch_bounds__(256) k_wgrad(WgradArgs w)
{
  const ConvArgs &a = w.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kq = lane >> 4;
  const long P = (long)a.B * a.D * a.H * a.W;
  const long p_begin = (long)blockIdx.x * w.chunk;
  const long p_end = p_begin + w.chunk < P ? p_begin + w.chunk : P;
  const int co = blockIdx.y * 16 + r16;     // A row
  const int kcol0 = (blockIdx.z * 4 + wave) * 16;
  const int kcol = kcol0 + r16;              // B column
  // decode the column's tap / channel once
  int tap = kcol / a.cin_pad;
  const int ci = kcol - tap * a.cin_pad;
  const bool kvalid = tap < (a.d3 ? 27 : 9) && kcol < a.K_pad;
  const int t3 = a.d3 ? tap : tap + 9;
  const int dz = t3 / 9 - 1, dy = (t3 % 9) / 3 - 1, dx = t3 % 3 - 1;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int W = a.W, H = a.H;
  for(long p0 = p_begin; p0 < p_end; p0 += 32) {
    bf16x8 av, bv;
    u16 *ah = (u16 *)&av, *bh = (u16 *)&bv;
    // coordinates of the lane's first voxel, then advanced one voxel at a time
    const long pb = p0 + 8 * kq;
    long q = pb;
    int x = (int)(q % W);
    q /= W;
    int y = (int)(q % H);
    q /= H;
    int z = (int)(q % a.D);
    int b = (int)(q / a.D);
#pragma unroll
    for(int j = 0; j < 8; j++) {
      const long p = pb + j;
      const bool ok = p < p_end;
      ah[j] = (ok && co < a.cout) ? w.g[p * a.cout + co] : (u16)0;
      float xv = 0.f;
      if(ok && kvalid) xv = load1<MODE>(a, b, z + dz, y + dy, x + dx, ci);
      bh[j] = f2bf(xv);
      if(++x == W) {
        x = 0;
        if(++y == H) {
          y = 0;
          if(++z == a.D) {
            z = 0;
            b++;
          }
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  // lane holds column kcol0 + r16, rows (co) blockIdx.y*16 + 4*kq + i
#pragma unroll
  for(int i = 0; i < 4; i++) {
    const int orow = blockIdx.y * 16 + 4 * kq + i;
    const int ocol = kcol0 + r16;
    if(orow < a.cout && ocol < a.K_pad)
      w.part[((long)blockIdx.x * a.cout + orow) * a.K_pad + ocol] = acc[i];
  }
}

// ---- weight gradient on channel-major copies (grids whose rows are multiples of 8)
//
// The MFMA reduction of the weight gradient runs over voxels, so its operands
// need 8 consecutive voxels of one channel per lane: the backward pass keeps
// channel-major copies GT[co][p] of the layer's output gradient and XT[ci][p]
// of its logical input (pooled / upsampled-and-concatenated as the forward
// loader forms it), written through an LDS transpose.  A voxel octet then is
// one aligned 16-byte load of GT and, for tap (dz, dy, dx), the octet of XT at
// p + (dz*H + dy)*W + dx: with W a multiple of 8 this is aligned for dx = 0
// and one element off for dx = +-1 (an aligned 16-byte load plus the
// neighbouring dword, realigned by byte shifts).  Octets whose neighbour row
// or plane lies outside the grid are zero; in-row x edges zero one element.
constexpr int kTrPix = 64;        // voxels per transpose block
constexpr int kTrStride = 72;     // LDS row stride (elements, 16-byte aligned rows)
constexpr int kXtPad = 16;        // elements before and after each XT copy

// G = BF16(dY * (Y > 0)) in both layouts: g[p][c] and gt[c][P] (P % 8 == 0)
__global__ void __launch_bounds__(256) k_relu_grad_t(long P, int C, const float *dy, const u16 *y, u16 *g,
                                                     u16 *gt)
{
  __shared__ u16 tile[128 * kTrStride];
  const long p0 = (long)blockIdx.x * kTrPix;
  for(int i = threadIdx.x; i < kTrPix * C; i += 256) {
    const int pl = i / C, c = i - pl * C;
    const long p = p0 + pl;
    u16 v = 0;
    if(p < P) {
      const long k = p * C + c;
      v = f2bf(bf2f(y[k]) > 0.f ? dy[k] : 0.f);
      g[k] = v;
    }
    tile[c * kTrStride + pl] = v;
  }
  __syncthreads();
  for(int i = threadIdx.x; i < C * (kTrPix / 8); i += 256) {
    const int c = i >> 3, seg = i & 7;
    const long p = p0 + 8 * seg;
    if(p < P) *(uint4 *)(gt + (long)c * P + p) = *(const uint4 *)(tile + c * kTrStride + 8 * seg);
  }
}

// XT[ci][p] = logical input of the convolution (the forward loader's octets)
template <int MODE>
__global__ void __launch_bounds__(256) k_input_t(ConvArgs a, u16 *xt)
{
  __shared__ u16 tile[192 * kTrStride];
  const long P = (long)a.B * a.D * a.H * a.W;
  const long p0 = (long)blockIdx.x * kTrPix;
  const int cin = a.cin_pad, noct = cin >> 3;
  for(int i = threadIdx.x; i < kTrPix * noct; i += 256) {
    const int pl = i % kTrPix, c8 = i / kTrPix;
    const long p = p0 + pl;
    uint4 v = {0u, 0u, 0u, 0u};
    if(p < P) {
      long q = p;
      const int px = (int)(q % a.W);
      q /= a.W;
      const int py = (int)(q % a.H);
      q /= a.H;
      const int pz = (int)(q % a.D);
      const int b = (int)(q / a.D);
      const long pix = MODE == M_POOL ? (a.d3 ? (((long)b * 2 * a.D + 2 * pz) * (2 * a.H) + 2 * py) * (2 * a.W) + 2 * px
                                              : ((long)b * 2 * a.H + 2 * py) * (2 * a.W) + 2 * px)
                                      : p;
      v = conv_load_a<MODE>(a, pix, pz, py, px, a.d3 ? b * (a.D >> 1) : b, 0, 0, 0, 8 * c8, true);
    }
    const u16 *e = (const u16 *)&v;
#pragma unroll
    for(int j = 0; j < 8; j++) tile[(8 * c8 + j) * kTrStride + pl] = e[j];
  }
  __syncthreads();
  for(int i = threadIdx.x; i < cin * (kTrPix / 8); i += 256) {
    const int c = i >> 3, seg = i & 7;
    const long p = p0 + 8 * seg;
    if(p < P) *(uint4 *)(xt + (long)c * P + p) = *(const uint4 *)(tile + c * kTrStride + 8 * seg);
  }
}

struct WgradTArgs {
  int D, H, W, d3, cin, cout, K_pad;
  long P;
  const u16 *gt;  // [cout][P]
  const u16 *xt;  // [cin][P], kXtPad readable elements before and after
  float *part;    // [nchunks][cout][K_pad]
  int chunk;      // voxels per chunk (multiple of 128)
  int nchunks, nco, nk;  // chunks, 16-row tiles of co, 64-column tiles of k
};

// XT octet of channel row xrow at voxel pb shifted by the tap (dz, dy, dx);
// (x0, y, z): coordinates of pb.  Zero outside the grid.
__device__ inline uint4 wgrad_b_octet(const u16 *xrow, long pb, long off, int dz, int dy, int dx, int x0, int y,
                                      int z, const WgradTArgs &w)
{
  uint4 bv = {0u, 0u, 0u, 0u};
  if((unsigned)(y + dy) >= (unsigned)w.H || (unsigned)(z + dz) >= (unsigned)w.D) return bv;
  const u16 *src = xrow + pb + off;
  if(dx == 0) return *(const uint4 *)src;
  if(dx > 0) {  // elements 1..8 of the aligned octet at src - 1
    const uint4 v = *(const uint4 *)(src - 1);
    const unsigned e = *(const unsigned *)(src + 7);
    bv.x = __builtin_amdgcn_alignbyte(v.y, v.x, 2);
    bv.y = __builtin_amdgcn_alignbyte(v.z, v.y, 2);
    bv.z = __builtin_amdgcn_alignbyte(v.w, v.z, 2);
    bv.w = __builtin_amdgcn_alignbyte(e, v.w, 2);
    if(x0 + 8 == w.W) bv.w &= 0xffffu;  // x + 1 leaves the row
    return bv;
  }
  // elements -1..6 of the aligned octet at src + 1
  const uint4 v = *(const uint4 *)(src + 1);
  const unsigned e = *(const unsigned *)(src - 1);
  bv.x = __builtin_amdgcn_alignbyte(v.x, e, 2);
  bv.y = __builtin_amdgcn_alignbyte(v.y, v.x, 2);
  bv.z = __builtin_amdgcn_alignbyte(v.z, v.y, 2);
  bv.w = __builtin_amdgcn_alignbyte(v.w, v.z, 2);
  if(x0 == 0) bv.x &= 0xffff0000u;  // x - 1 leaves the row
  return bv;
}

__device__ inline void wgrad_step(int &x, int &y, int &z, int n, const WgradTArgs &w)
{
  x += n;
  while(x >= w.W) {
    x -= w.W;
    if(++y == w.H) {
      y = 0;
      if(++z == w.D) z = 0;
    }
  }
}

// dW[co][k] over one chunk of voxels: wave tile 16 co x 16 columns k = (tap, ci).
// The MFMA sum over voxels is order-free, so lane group q takes the 32
// consecutive voxels p0 + 32 q .. + 31 as four octets of four MFMA steps (64
// contiguous bytes of each operand row per lane).  Workgroups are ordered so
// that all tiles of one chunk run on one XCD (workgroup i runs on XCD i % 8),
// which keeps the chunk's GT and XT rows in that XCD's L2.
__global__ void __launch_bounds__(256) k_wgrad_t(WgradTArgs w)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, q = lane >> 4;
  const int T = w.nco * w.nk;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int chunk_id = xcd + 8 * (slot / T);
  if(chunk_id >= w.nchunks) return;
  const int tile = slot % T;
  const int ty = tile / w.nk, tz = tile - ty * w.nk;
  const long p_begin = (long)chunk_id * w.chunk;
  const long p_end = p_begin + w.chunk < w.P ? p_begin + w.chunk : w.P;
  const int co = ty * 16 + r16;
  const int kcol = (tz * 4 + wave) * 16 + r16;
  const int tap = kcol / w.cin;
  const int ci = kcol - tap * w.cin;
  const bool kvalid = tap < (w.d3 ? 27 : 9);
  const int t3 = w.d3 ? tap : tap + 9;
  const int dz = t3 / 9 - 1, dy = (t3 % 9) / 3 - 1, dx = t3 % 3 - 1;
  const long off = ((long)dz * w.H + dy) * w.W + dx;
  const u16 *arow = w.gt + (long)(co < w.cout ? co : 0) * w.P;
  const u16 *xrow = w.xt + (long)(kvalid ? ci : 0) * w.P;
  const bool aok = co < w.cout;
  long pb = p_begin + 32 * q;
  int x0, y, z;
  {
    long t = pb;
    x0 = (int)(t % w.W);
    t /= w.W;
    y = (int)(t % w.H);
    t /= w.H;
    z = (int)(t % w.D);
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  for(long p0 = p_begin; p0 < p_end; p0 += 128, pb += 128) {
    uint4 av[4], bv[4];
    int xj = x0, yj = y, zj = z;
#pragma unroll
    for(int j = 0; j < 4; j++) {
      const long pj = pb + 8 * j;
      const bool in = pj < p_end;
      av[j] = (in && aok) ? *(const uint4 *)(arow + pj) : zero4;
      bv[j] = (in && kvalid) ? wgrad_b_octet(xrow, pj, off, dz, dy, dx, xj, yj, zj, w) : zero4;
      wgrad_step(xj, yj, zj, 8, w);
    }
#pragma unroll
    for(int j = 0; j < 4; j++)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[j]), as_bf8(bv[j]), acc, 0, 0, 0);
    wgrad_step(x0, y, z, 128, w);
  }
#pragma unroll
  for(int i = 0; i < 4; i++) {
    const int orow = ty * 16 + 4 * q + i;
    const int ocol = (tz * 4 + wave) * 16 + r16;
    if(orow < w.cout && ocol < w.K_pad) w.part[((long)chunk_id * w.cout + orow) * w.K_pad + ocol] = acc[i];
  }
}

// Weight gradient on LDS tiles (grids the convolution tiles divide): a
// workgroup takes a run of `ntile` output tiles (16 x 16 pixels or 4 x 4 x 16
// voxels), 16 output channels co and CC input channels ci.  Per tile, G^T
// (the tile's voxels x 16 co) and the input halo (CC channels, channel-major,
// halo rows padded to 32 entries with x = -1 at entry 7) are staged in LDS;
// the MFMA reduction runs over the tile's voxels, 32 (two tile rows) per step,
// wave w taking taps w, w + 4, ...: the G^T fragment is shared by the wave's
// taps, the input fragment of tap (dz, dy, dx) is the halo row segment at
// x + dx (one element off for dx = +-1: an aligned 16-byte read plus the
// neighbouring dword, realigned by byte shifts).
constexpr int kWgRow = 32;  // padded halo row (entries)
struct WgradLArgs {
  ConvArgs x;        // logical input (MODE loader) and sizes at this level
  const u16 *gt;     // [cout][P]
  float *part;       // [nchunks][cout][K_pad]
  int CC, ntile, ntiles;
};

template <int MODE, int CCN>  // CCN = CC / 16 input-channel subtiles
__global__ void __launch_bounds__(256) k_wgrad_lds(WgradLArgs w)
{
  extern __shared__ uint4 wl[];
  const ConvArgs &a = w.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kq = lane >> 4;
  const int d3 = a.d3, TY = tile_ty(d3), TZ = tile_tz(d3);
  const int HY = TY + 2, HZ = d3 ? TZ + 2 : 1, nrow = HZ * HY;
  const int nx = a.W / kTileX, ny = a.H / TY, nz = d3 ? a.D / TZ : 1;
  const int taps = d3 ? 27 : 9, t0 = d3 ? 0 : 9;
  const int CC = 16 * CCN, cin = a.cin_pad, cout = a.cout;
  const int co0 = blockIdx.y * 16, c0 = blockIdx.z * CC;
  const long P = (long)a.B * a.D * a.H * a.W;
  u16 *gs = (u16 *)wl;                 // [16 co][256 voxels]
  u16 *xs = gs + 16 * 256;             // [CC][nrow][kWgRow]
  f32x4 acc[7][CCN];
#pragma unroll
  for(int i = 0; i < 7; i++)
#pragma unroll
    for(int c = 0; c < CCN; c++) acc[i][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const uint4 zero4 = {0u, 0u, 0u, 0u};
  const int tile0 = blockIdx.x * w.ntile;
  for(int ti = 0; ti < w.ntile; ti++) {
    const long tile = tile0 + ti;
    if(tile >= w.ntiles) break;
    long t = tile;
    const int tx = (int)(t % nx);
    t /= nx;
    const int ty = (int)(t % ny);
    t /= ny;
    const int tz = (int)(t % nz);
    const int b = (int)(t / nz);
    const int x0 = tx * kTileX, y0 = ty * TY, z0 = tz * TZ;
    __syncthreads();
    // G^T of the tile: 16 co x 16 rows x 16 x, one 16-byte octet per thread pair
    for(int i = threadIdx.x; i < 16 * 16 * 2; i += 256) {
      const int co = i >> 5, rr = (i >> 1) & 15, h = i & 1;
      const int zz = d3 ? rr / TY : 0, yy = d3 ? rr % TY : rr;
      const long p = (((long)b * a.D + z0 + zz) * a.H + y0 + yy) * a.W + x0 + 8 * h;
      uint4 v = zero4;
      if(co0 + co < cout) v = *(const uint4 *)(w.gt + (long)(co0 + co) * P + p);
      *(uint4 *)(gs + co * 256 + rr * 16 + 8 * h) = v;
    }
    // input halo, channel-major: halo voxel hx (x = x0 - 1 + hx) at entry 7 + hx
    for(int i = threadIdx.x; i < nrow * 18 * CCN * 2; i += 256) {
      const int g = i % (CCN * 2), v = i / (CCN * 2);
      const int hx = v % 18, hr = v / 18;
      const int hy = hr % HY, hz = hr / HY;
      const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = d3 ? z0 - 1 + hz : 0;
      uint4 o = zero4;
      if((unsigned)x < (unsigned)a.W && (unsigned)y < (unsigned)a.H && (unsigned)z < (unsigned)a.D &&
         c0 + 8 * g < cin) {  // channels past cin (cin = 8) stay zero
        const long pix = MODE == M_POOL ? (d3 ? (((long)b * 2 * a.D + 2 * z) * (2 * a.H) + 2 * y) * (2 * a.W) + 2 * x
                                              : ((long)b * 2 * a.H + 2 * y) * (2 * a.W) + 2 * x)
                                        : (((long)b * a.D + z) * a.H + y) * a.W + x;
        o = conv_load_a<MODE>(a, pix, z, y, x, d3 ? b * (a.D >> 1) : b, 0, 0, 0, c0 + 8 * g, true);
      }
      const u16 *e = (const u16 *)&o;
#pragma unroll
      for(int j = 0; j < 8; j++) xs[((8 * g + j) * nrow + hr) * kWgRow + 7 + hx] = e[j];
    }
    // halo row padding entries (outside x = -1 .. 16) stay unread
    __syncthreads();
    for(int s = 0; s < 8; s++) {  // 32 voxels: tile rows 2s, 2s + 1
      const int rr = 2 * s + (kq >> 1), xo = 8 * (kq & 1);
      const uint4 av = *(const uint4 *)(gs + r16 * 256 + rr * 16 + xo);
      const int zz = d3 ? rr / TY : 0, yy = d3 ? rr % TY : rr;
#pragma unroll
      for(int i = 0; i < 7; i++) {
        const int tap = wave + 4 * i;
        if(tap >= taps) break;
        const int tt = tap + t0;
        const int dz = tap_dz(tt), dy = tap_dy(tt), dx = tap_dx(tt);
        const int hr = (zz + (d3 ? 1 + dz : 0)) * HY + (yy + 1 + dy);
#pragma unroll
        for(int c = 0; c < CCN; c++) {
          // entries x = xo + dx .. + 7 of halo row hr, channel 16 c + r16 (x = -1 at entry 7)
          const u16 *row = xs + ((16 * c + r16) * nrow + hr) * kWgRow + 8 + xo;
          uint4 bv;
          if(dx == 0) bv = *(const uint4 *)row;
          else if(dx > 0) {
            const uint4 v = *(const uint4 *)row;
            const unsigned e = *(const unsigned *)(row + 8);
            bv.x = __builtin_amdgcn_alignbyte(v.y, v.x, 2);
            bv.y = __builtin_amdgcn_alignbyte(v.z, v.y, 2);
            bv.z = __builtin_amdgcn_alignbyte(v.w, v.z, 2);
            bv.w = __builtin_amdgcn_alignbyte(e, v.w, 2);
          }
          else {
            const uint4 v = *(const uint4 *)row;
            const unsigned e = *(const unsigned *)(row - 2);
            bv.x = __builtin_amdgcn_alignbyte(v.x, e, 2);
            bv.y = __builtin_amdgcn_alignbyte(v.y, v.x, 2);
            bv.z = __builtin_amdgcn_alignbyte(v.z, v.y, 2);
            bv.w = __builtin_amdgcn_alignbyte(v.w, v.z, 2);
          }
          acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av), as_bf8(bv), acc[i][c], 0, 0, 0);
        }
      }
    }
  }
  // partials: rows co0 + 4 kq + j, columns tap * cin + c0 + 16 c + r16
  float *out = w.part + (long)blockIdx.x * cout * a.K_pad;
#pragma unroll
  for(int i = 0; i < 7; i++) {
    const int tap = wave + 4 * i;
    if(tap >= taps) break;
#pragma unroll
    for(int c = 0; c < CCN; c++)
#pragma unroll
      for(int j = 0; j < 4; j++) {
        const int co = co0 + 4 * kq + j, ci = c0 + 16 * c + r16;
        if(co < cout && ci < cin) out[(long)co * a.K_pad + tap * cin + ci] = acc[i][c][j];
      }
  }
  // the padding columns of the weight layout
  if(blockIdx.z == 0)
    for(int i = threadIdx.x; i < 16 * (a.K_pad - taps * cin); i += 256) {
      const int co = co0 + i / (a.K_pad - taps * cin), k = taps * cin + i % (a.K_pad - taps * cin);
      if(co < cout) out[(long)co * a.K_pad + k] = 0.f;
    }
}

// out[i] = sum over chunks of part[c][i]: a block takes 32 elements, its 8
// thread groups the chunks g, g + 8, ..., then the 8 group sums are added in
// group order
__global__ void __launch_bounds__(256) k_sum_chunks(int nchunks, long n, const float *part, float *out)
{
  __shared__ float sh[8][32];
  const int e = threadIdx.x & 31, g = threadIdx.x >> 5;
  const long i = (long)blockIdx.x * 32 + e;
  float s = 0.f;
  if(i < n)
    for(int c = g; c < nchunks; c += 8) s += part[(long)c * n + i];
  sh[g][e] = s;
  __syncthreads();
  if(g == 0 && i < n) {
    float t = 0.f;
#pragma unroll
    for(int k = 0; k < 8; k++) t += sh[k][e];
    out[i] = t;
  }
}

// G = BF16(dY * (Y > 0)), with Y the layer's (ReLU) output
__global__ void k_relu_grad(long n, const float *dy, const u16 *y, u16 *g)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i < n) g[i] = f2bf(bf2f(y[i]) > 0.f ? dy[i] : 0.f);
}

// bias gradient: db[c] = sum_p G[p][c]  (one block per channel)
// Column sums of a [P][C] tensor (C divides 256), in two deterministic stages.
// Stage 1: block b takes elements [b * 8192, (b + 1) * 8192); thread t reads
// elements t, t + 256, ... (coalesced), all in column t % C, optionally
// weighted by dz[pixel]; the threads of one column are then added in thread
// order.  Stage 2 adds the blocks' partials in block order.
constexpr int kColChunk = 8192;
template <typename T>
__global__ void __launch_bounds__(256) k_colsum_part(long n, int C, const T *x, const float *dz, float *part)
{
  __shared__ float sh[256];
  const int t = threadIdx.x;
  const long b0 = (long)blockIdx.x * kColChunk;
  float acc = 0.f;
  for(int k = 0; k < kColChunk / 256; k++) {
    const long i = b0 + t + 256L * k;
    if(i >= n) break;
    float v;
    if constexpr(sizeof(T) == 2) v = bf2f(x[i]);
    else v = x[i];
    if(dz) v *= dz[i / C];
    acc += v;
  }
  sh[t] = acc;
  __syncthreads();
  if(t < C) {
    float s = 0.f;
    for(int j = t; j < 256; j += C) s += sh[j];
    part[(long)blockIdx.x * C + t] = s;
  }
}

// one block per column: thread t adds blocks t, t + 256, ... in order, then a
// fixed-order tree over the threads
__global__ void __launch_bounds__(256) k_colsum_fin(int nblocks, int C, const float *part, float *out)
{
  const int c = blockIdx.x;
  float s = 0.f;
  for(int b = threadIdx.x; b < nblocks; b += 256) s += part[(long)b * C + c];
  s = block_sum<float, 256>(s);
  if(threadIdx.x == 0) out[c] = s;
}

// adjoints of the fused loaders, accumulating into FP32 gradient tensors
__global__ void k_add(long n, const float *src, float *dst)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i < n) dst[i] += src[i];
}

// pool^T: dFine[2z+c][2y+a][2x+b][ch] += dX[z][y][x][ch] / 8 (3D), or
// dFine[2y+a][2x+b][ch] += dX[y][x][ch] / 4 (2D, D = 1)
__global__ void k_pool_adj(int B, int D, int H, int W, int C, int d3, const float *dx, float *dfine)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  const int Df = d3 ? 2 * D : 1;
  long n = (long)B * Df * 2 * H * 2 * W * C;
  if(i >= n) return;
  int c = (int)(i % C);
  long q = i / C;
  int xf = (int)(q % (2 * W));
  q /= 2 * W;
  int yf = (int)(q % (2 * H));
  q /= 2 * H;
  int zf = (int)(q % Df);
  int b = (int)(q / Df);
  const int zc = d3 ? zf >> 1 : 0;
  dfine[i] += (d3 ? 0.125f : 0.25f) * dx[((((long)b * D + zc) * H + (yf >> 1)) * W + (xf >> 1)) * C + c];
}

// up^T (++ split): dCoarse[zc][yc][xc][c] += sum of the 8 (2D: 4) fine
// dX[.][.][.][c]; dSkip[z][y][x][c] += dX[z][y][x][c1 + c]
__global__ void k_upcat_adj(int B, int D, int H, int W, int d3, int c1, int c2, const float *dx, float *dcoarse,
                            float *dskip)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  const int C = c1 + c2;
  const int Dc = d3 ? D / 2 : 1;
  long ncoarse = (long)B * Dc * (H / 2) * (W / 2) * c1;
  long nskip = (long)B * D * H * W * c2;
  if(i < ncoarse) {
    int c = (int)(i % c1);
    long q = i / c1;
    int xc = (int)(q % (W / 2));
    q /= W / 2;
    int yc = (int)(q % (H / 2));
    q /= H / 2;
    int zc = (int)(q % Dc);
    int b = (int)(q / Dc);
    float s = 0.f;
    for(int c2z = 0; c2z < (d3 ? 2 : 1); c2z++)
      for(int a2 = 0; a2 < 2; a2++)
        for(int b2 = 0; b2 < 2; b2++)
          s += dx[((((long)b * D + (d3 ? 2 * zc + c2z : 0)) * H + 2 * yc + a2) * W + 2 * xc + b2) * C + c];
    dcoarse[i] += s;
  }
  else if(i < ncoarse + nskip) {
    long j = i - ncoarse;
    int c = (int)(j % c2);
    long pix = j / c2;
    dskip[j] += dx[pix * C + c1 + c];
  }
}

// 1x1 output convolution (16 -> 1), FP32, at pixel p
__device__ inline float out_pixel(const u16 *x, long p, int C, const float *w, const float *b)
{
  float s = b[0];
  for(int c = 0; c < C; c++) s += bf2f(x[p * C + c]) * w[c];
  return s;
}

__global__ void k_out_fwd(long P, int C, const u16 *x, const float *w, const float *b, float *z,
                          const int *gate, int gmode)
{
  if(unet_gated(gate, gmode)) return;
  long p = (long)blockIdx.x * 256 + threadIdx.x;
  if(p >= P) return;
  z[p] = out_pixel(x, p, C, w, b);
}

__global__ void k_out_bwd_x(long P, int C, const float *dz, const float *w, float *dx)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i >= P * C) return;
  dx[i] += dz[i / C] * w[i % C];
}

// gw[c] = sum_p dz[p] x[p][c] (c < C), gb = sum_p dz[p] (c == C); one block each

// scale-invariant loss per sample: L = | z/|z| - t |^2 over the mask, t unit
// norm; dz = -(2/|z|) (t - (zhat.t) zhat).  One block per sample.
__global__ void k_loss(int HW, const float *z, const float *t, const float *mask, float *dz,
                       float *loss)
{
  const int s = blockIdx.x;
  const float *zs = z + (long)s * HW, *ts = t + (long)s * HW, *ms = mask;
  float zz = 0.f, zt = 0.f;
  for(int i = threadIdx.x; i < HW; i += 256) {
    float m = ms[i], zv = zs[i] * m;
    zz += zv * zv;
    zt += zv * ts[i];
  }
  __shared__ float sh[2];
  float a = block_sum<float, 256>(zz);
  if(threadIdx.x == 0) sh[0] = a;
  __syncthreads();
  float b = block_sum<float, 256>(zt);
  if(threadIdx.x == 0) sh[1] = b;
  __syncthreads();
  const float nz = sqrtf(fmaxf(sh[0], 1e-30f));
  const float c = sh[1] / nz;  // zhat . t
  if(threadIdx.x == 0) loss[s] = 2.f - 2.f * c;
  for(int i = threadIdx.x; i < HW; i += 256) {
    float m = ms[i];
    float zh = zs[i] * m / nz;
    dz[(long)s * HW + i] = -(2.f / nz) * (ts[i] - c * zh) * m;
  }
}

// Adam (FP32 master weights)
__global__ void k_adam(long n, float *w, const float *g, float *m, float *v, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, float wd)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if(i >= n) return;
  float gi = g[i] + wd * w[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 forward weights and the tap-mirrored, transposed adjoint weights
__global__ void k_prep_weights(int cout, int cin_pad, int K_pad, int K2_pad, int taps, const float *w32,
                               u16 *wt, u16 *wtT)
{
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  long n = (long)cout * K_pad;
  if(i < n) wt[i] = f2bf(w32[i]);
  long n2 = (long)cin_pad * K2_pad;
  if(i < n2) {
    int ci = (int)(i / K2_pad), k2 = (int)(i % K2_pad);
    int tap2 = k2 / cout, co = k2 - tap2 * cout;
    float v = 0.f;
    if(tap2 < taps) v = w32[(long)co * K_pad + (taps - 1 - tap2) * cin_pad + ci];
    wtT[i] = f2bf(v);
  }
}

// network input: ch0 = residual (normalised), ch1 = coefficient feature, rest 0
__global__ void k_make_input(long P, const float *r, const float *feat, u16 *x)
{
  long p = (long)blockIdx.x * 256 + threadIdx.x;
  if(p >= P) return;
  uint4 v = {0u, 0u, 0u, 0u};
  u16 *h = (u16 *)&v;
  h[0] = f2bf(r[p]);
  h[1] = f2bf(feat[p]);
  *(uint4 *)(x + p * 8) = v;
}

// ---- preconditioner glue (solver vectors <-> grid) ----
// The solver works on the renumbered, diagonally scaled system
//   Ah = D P A P^T D,  rh = D P r,  x = P^T D xh      (D = rs, P = perm)
// so Ah^{-1} rh = D^{-1} P A^{-1} r: unscale/unpermute the input, run the
// network on the original grid, then permute/unscale the output.
// perm: original -> internal index, rperm: internal -> original, rs indexed
// by original DOF (null: no scaling).
template <typename TV>
__global__ void __launch_bounds__(256) k_pc_unscale(int npix, const int *pix2dof, const int *perm,
                                                    const double *rs, const double *dscale,
                                                    const TV *in, float *r32, double *part,
                                                    const int *gate, int gmode)
{
  if(unet_gated(gate, gmode)) return;
  double s = 0.;
  for(long p = (long)blockIdx.x * 256 + threadIdx.x; p < npix; p += (long)gridDim.x * 256) {
    const int d = pix2dof[p];
    double v = 0.;
    if(d >= 0) {
      v = (double)in[perm ? perm[d] : d];
      if(rs) v /= rs[d];
      if(dscale) v *= dscale[d];  // network space: D^{-1/2} r
    }
    r32[p] = (float)v;
    s += v * v;
  }
  s = block_sum<double, 256>(s);
  if(threadIdx.x == 0) part[blockIdx.x] = s;
}

// network input: ch0 = r * sqrt(ndof)/|r| (unit RMS over the DOFs), ch1 = feature
__global__ void __launch_bounds__(256) k_pc_gather(int npix, int ndof, const float *r32,
                                                   const double *part, int nparts, const u16 *feat,
                                                   u16 *x, int odd, const int *gate, int gmode)
{
  if(unet_gated(gate, gmode)) return;
  __shared__ double s_inv;
  double t = sum_partials<double, 256>(part, nparts);
  if(threadIdx.x == 0) s_inv = t > 0. ? sqrt((double)ndof / t) : 0.;
  __syncthreads();
  const float inv = (float)s_inv;
  for(long p = (long)blockIdx.x * 256 + threadIdx.x; p < npix; p += (long)gridDim.x * 256) {
    uint4 o = {0u, 0u, 0u, 0u};
    u16 *h = (u16 *)&o;
    h[0] = f2bf(r32[p] * inv);
    h[1] = feat[p];
    *(uint4 *)(x + p * 8) = o;
    if(odd) {  // second sample: the negated residual
      h[0] ^= 0x8000;
      *(uint4 *)(x + (npix + p) * 8) = o;
    }
  }
}

// The network's output layer (the 1x1 convolution of the last activation x,
// evaluated here at the DOF's pixel) mapped back to the solver's vector.
template <typename TV>
__global__ void __launch_bounds__(256) k_pc_scatter(int n, int ndof, const int *rperm,
                                                    const int *dof2pix, const double *rs,
                                                    const double *dscale, const u16 *x,
                                                    const float *ow, const float *ob,
                                                    const double *part, int nparts,
                                                    int npix, int odd, TV *out, const int *gate,
                                                    int gmode)
{
  if(unet_gated(gate, gmode)) return;
  __shared__ double s_sc;
  double t = sum_partials<double, 256>(part, nparts);
  if(threadIdx.x == 0) s_sc = sqrt(t / (double)ndof);
  __syncthreads();
  const double sc = s_sc;
  for(long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int d = rperm ? rperm[i] : (int)i;
    const int px = dof2pix[d];
    // odd part N(r) - N(-r) removes what does not depend on the residual
    const float z0 = out_pixel(x, px, 16, ow, ob);
    double v = sc * (odd ? 0.5 * ((double)z0 - (double)out_pixel(x, (long)npix + px, 16, ow, ob))
                         : (double)z0);
    if(dscale) v *= dscale[d];  // back from D^{1/2} e
    if(rs) v /= rs[d];
    out[i] = (TV)v;
  }
}

long rup(long a, long b) { return (a + b - 1) / b * b; }

}  // namespace

// =============================================================================
// Network object
// =============================================================================
struct ConvLayer {
  int mode = M_PLAIN, cin = 0, cin_pad = 0, c1 = 0, cout = 0, K_pad = 0, K2_pad = 0;
  int in_tensor = -1, in2_tensor = -1, level = 0;  // activation indices; level: spatial /2^level
  float *w32 = nullptr, *b32 = nullptr, *gw = nullptr, *gb = nullptr;
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
  u16 *wt = nullptr, *wtT = nullptr;
};

struct UserPrec {
  int H = 0, W = 0, Bmax = 0;  // grid and batch capacity
  int D = 1, d3 = 0, taps = 9; // depth (1: 2D grid), 3D flag, taps per convolution
  // voxels of one sample at level lvl (spatial size / 2^lvl in each direction)
  long vox(int lvl) const { return (long)(d3 ? D >> lvl : 1) * (H >> lvl) * (W >> lvl); }
  std::vector<ConvLayer> L;    // 14 convolutions
  float *ow = nullptr, *ob = nullptr, *gow = nullptr, *gob = nullptr;  // 1x1 output (16 + 1)
  float *mow = nullptr, *vow = nullptr, *mob = nullptr, *vob = nullptr;
  // activations (BF16) and their gradients (FP32), one per layer output + input
  std::vector<u16 *> act;    // [0] = input X0 (8 ch), [i+1] = output of conv i
  std::vector<float *> gact; // gradients of act
  std::vector<int> act_ch, act_lvl;
  float *z = nullptr, *dz = nullptr, *target = nullptr, *mask = nullptr, *loss = nullptr;
  float *dx = nullptr;  // scratch for the adjoint of one convolution's input
  u16 *g = nullptr;     // ReLU-masked gradient (BF16) of one layer
  u16 *gt = nullptr;    // the same, channel-major [cout][P]
  u16 *xt = nullptr;    // logical input of one layer, channel-major [cin][P] (padded)
  float *part = nullptr;
  size_t part_size = 0;
  float *cpart = nullptr;  // column-sum partials (bias and output-layer gradients)
  size_t cpart_size = 0;
  float *split = nullptr;  // split-K partials of the convolutions
  size_t split_size = 0;
  int adam_t = 0;
  // preconditioner binding
  int n_dof = 0;
  int *pix2dof = nullptr, *dof2pix = nullptr;
  u16 *feat = nullptr;
  float *r32 = nullptr;
  double *dscale = nullptr;  // per original DOF, null: the network works on unscaled vectors
  double *pc_part = nullptr;
  int pc_nparts = 0;
  int odd = 0;  // apply the odd part of the network (needs batch >= 2)
  int lsq = 0;  // two-stage: scale the network output by its least-squares factor
  // the solver's renumbering and scaling of the current solve
  const int *perm = nullptr, *rperm = nullptr;
  const double *rs = nullptr;
  ~UserPrec();
};

namespace {

template <typename T> void zfree(T *&p) { dfree(p); }

ConvArgs conv_args(const UserPrec &U, const ConvLayer &l, int B)
{
  ConvArgs a{};
  a.B = B;
  a.H = U.H >> l.level;
  a.W = U.W >> l.level;
  a.D = U.d3 ? U.D >> l.level : 1;
  a.d3 = U.d3;
  a.cin_pad = l.cin_pad;
  a.cout = l.cout;
  a.K_pad = l.K_pad;
  a.c1 = l.c1;
  a.in = U.act[l.in_tensor];
  a.in2 = l.in2_tensor >= 0 ? U.act[l.in2_tensor] : nullptr;
  a.wt = l.wt;
  a.bias = l.b32;
  a.relu = 1;
  return a;
}

// Launch shape of one convolution: channel tile TN from cout, pixel tile TM
// (64 or 16 pixels per wave) and split-K so that a launch has ~1024 waves
// (4 per SIMD of the 256 CUs) even on the coarse levels of a single grid.
struct ConvPlan {
  int TM, TN, nsplit, kchunk;
  bool ks;            // 4 waves per tile, reduction split inside the block
  long split_floats;  // partial-sum buffer needed (0 without split-K)
};

// Launch shape of one convolution: about one wave per SIMD (1024 on 256 CUs;
// measured at batch 1, 256^2: 4 or 16 waves per SIMD are no faster, the
// per-layer time is the same for every shape tried): 64 pixels per wave when
// that fills the GPU, else 16; for smaller images, a 16-pixel tile shared by
// the 4 waves of a workgroup (reduction split 4 ways), with narrower column
// tiles until the grid has the target number of waves; for short reductions,
// split-K over the grid.
ConvPlan plan_conv(long P, int cout, int K_pad)
{
  ConvPlan c;
  // GDP_CONV_WAVES overrides the wave target (launch-shape sweeps)
  static long target = -1;
  if(target < 0) {
    const char *e = std::getenv("GDP_CONV_WAVES");
    target = (e && std::atol(e) > 0) ? std::atol(e) : 1024;
  }
  auto ntn = [&](int tn) { return (long)(cout + 16 * tn - 1) / (16 * tn); };
  c.TN = cout <= 16 ? 1 : (cout <= 32 ? 2 : 4);
  c.TM = 4;
  c.ks = false;
  long waves = (P + 63) / 64 * ntn(c.TN);
  if(waves < target) {
    c.TM = 1;
    waves = (P + 15) / 16 * ntn(c.TN);
  }
  if(waves < target && K_pad >= 4 * 32) {
    // share each tile between 4 waves; no partial buffer, no extra kernel
    c.ks = true;
    while(c.TN > 1 && 4 * ((P + 15) / 16) * ntn(c.TN) < target) c.TN >>= 1;
    c.nsplit = 1;
    c.kchunk = K_pad;
    c.split_floats = 0;
    return c;
  }
  const int ksteps = K_pad / 32;
  c.nsplit = 1;
  while(waves * c.nsplit < target && c.nsplit * 4 <= ksteps) c.nsplit *= 2;
  c.kchunk = ((ksteps + c.nsplit - 1) / c.nsplit) * 32;
  c.nsplit = (K_pad + c.kchunk - 1) / c.kchunk;
  c.split_floats = c.nsplit > 1 ? (long)c.nsplit * P * cout : 0;
  return c;
}

template <int MODE, int TM> int launch_conv_tm(const ConvArgs &a, int TN, dim3 grid)
{
  hipStream_t s = ctx().stream;
  if(TN == 1) hipLaunchKernelGGL((k_conv<MODE, TM, 1>), grid, dim3(256), 0, s, a);
  else if(TN == 2) hipLaunchKernelGGL((k_conv<MODE, TM, 2>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_conv<MODE, TM, 4>), grid, dim3(256), 0, s, a);
  return GDP_OK;
}

template <int MODE> int launch_conv_mode(ConvArgs a, const ConvPlan &c, float *split)
{
  const long P = (long)a.B * a.D * a.H * a.W;
  a.kchunk = c.kchunk;
  a.split = c.nsplit > 1 ? split : nullptr;
  if(c.ks) {
    dim3 grid((unsigned)((P + 15) / 16), (unsigned)((a.cout + 16 * c.TN - 1) / (16 * c.TN)));
    hipStream_t s = ctx().stream;
    if(c.TN == 1) hipLaunchKernelGGL((k_conv_ks<MODE, 1>), grid, dim3(256), 0, s, a);
    else if(c.TN == 2) hipLaunchKernelGGL((k_conv_ks<MODE, 2>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_conv_ks<MODE, 4>), grid, dim3(256), 0, s, a);
    GDP_HIP(hipGetLastError());
    return GDP_OK;
  }
  const long ptile = 4L * 16 * c.TM;
  dim3 grid((unsigned)((P + ptile - 1) / ptile), (unsigned)((a.cout + 16 * c.TN - 1) / (16 * c.TN)),
            (unsigned)c.nsplit);
  if(c.TM == 4) launch_conv_tm<MODE, 4>(a, c.TN, grid);
  else launch_conv_tm<MODE, 1>(a, c.TN, grid);
  if(c.nsplit > 1)
    hipLaunchKernelGGL(k_conv_split_sum, dim3((unsigned)((P * a.cout + 255) / 256)), dim3(256), 0,
                       ctx().stream, a, c.nsplit);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

dim3 g1(long n) { return dim3((unsigned)((n + 255) / 256)); }

// LDS-staged convolution: usable when the tiles divide the grid and give at
// least 256 workgroups; CC (channels staged at a time) is the largest
// multiple of 8 dividing cin whose halo fits kConvLds bytes
constexpr int kConvLds = 78 * 1024;
int conv_lds_cc(const ConvArgs &a)
{
  const int TY = tile_ty(a.d3), TZ = tile_tz(a.d3);
  if(a.W % kTileX || a.H % TY || (a.d3 && a.D % TZ)) return 0;
  const long tiles = (long)a.B * (a.W / kTileX) * (a.H / TY) * (a.d3 ? a.D / TZ : 1);
  if(tiles * ((a.cout + 15) / 16) < 128) return 0;
  const long nvox = (long)(kTileX + 2) * (TY + 2) * (a.d3 ? TZ + 2 : 1);
  for(int cc = a.cin_pad; cc >= 8; cc -= 8)
    if(a.cin_pad % cc == 0 && nvox * (cc / 8 + 1) * 16 <= kConvLds) return cc;
  return 0;
}

// dynamic LDS beyond the default limit, once per process (at network creation,
// never under graph capture)
template <int MODE> int conv_lds_attr_mode()
{
  GDP_HIP(hipFuncSetAttribute((const void *)k_conv_lds<MODE, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
  GDP_HIP(hipFuncSetAttribute((const void *)k_conv_lds<MODE, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
  GDP_HIP(hipFuncSetAttribute((const void *)k_conv_lds<MODE, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
  return GDP_OK;
}
template <int MODE> int wgrad_lds_attr_mode()
{
  GDP_HIP(hipFuncSetAttribute((const void *)k_wgrad_lds<MODE, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
  GDP_HIP(hipFuncSetAttribute((const void *)k_wgrad_lds<MODE, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
  return GDP_OK;
}
int conv_lds_attrs()
{
  static int rc = -1;
  if(rc < 0) {
    rc = conv_lds_attr_mode<M_PLAIN>();
    if(!rc) rc = conv_lds_attr_mode<M_POOL>();
    if(!rc) rc = conv_lds_attr_mode<M_UPCAT>();
    if(!rc) rc = wgrad_lds_attr_mode<M_PLAIN>();
    if(!rc) rc = wgrad_lds_attr_mode<M_POOL>();
    if(!rc) rc = wgrad_lds_attr_mode<M_UPCAT>();
  }
  return rc;
}

// LDS-tiled weight gradient: channels per workgroup (16 or 32, 0: not
// applicable -- tiles must divide the grid and CC divide cin, or cin = 8
// with the upper half of the 16 channels zero)
int wgrad_lds_cc(const ConvArgs &a)
{
  const int TY = tile_ty(a.d3), TZ = tile_tz(a.d3);
  if(a.W % kTileX || a.H % TY || (a.d3 && a.D % TZ) || (a.cin_pad % 16 && a.cin_pad != 8)) return 0;
  const long nrow = (long)(TY + 2) * (a.d3 ? TZ + 2 : 1);
  if(a.cin_pad % 32 == 0 && 16 * 256 * 2 + 32 * nrow * kWgRow * 2 <= kConvLds) return 32;
  return 16;
}

template <int MODE, int TN> int launch_conv_lds_tn(const ConvArgs &a, int cc)
{
  const int TY = tile_ty(a.d3), TZ = tile_tz(a.d3);
  const long tiles = (long)a.B * (a.W / kTileX) * (a.H / TY) * (a.d3 ? a.D / TZ : 1);
  const size_t lds = (size_t)(kTileX + 2) * (TY + 2) * (a.d3 ? TZ + 2 : 1) * (cc / 8 + 1) * 16;
  dim3 grid((unsigned)tiles, (unsigned)((a.cout + 16 * TN - 1) / (16 * TN)));
  hipLaunchKernelGGL((k_conv_lds<MODE, TN>), grid, dim3(256), lds, ctx().stream, a, cc);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

// channel tile: the widest (most reuse of each staged input fragment) that
// still gives 256 workgroups
template <int MODE> int launch_conv_lds(const ConvArgs &a, int cc)
{
  const long tiles = (long)a.B * (a.W / kTileX) * (a.H / tile_ty(a.d3)) * (a.d3 ? a.D / tile_tz(a.d3) : 1);
  auto wgs = [&](int tn) { return tiles * ((a.cout + 16 * tn - 1) / (16 * tn)); };
  if(a.cout > 32 && wgs(4) >= 256) return launch_conv_lds_tn<MODE, 4>(a, cc);
  if(a.cout > 16 && wgs(2) >= 256) return launch_conv_lds_tn<MODE, 2>(a, cc);
  return launch_conv_lds_tn<MODE, 1>(a, cc);
}

// Plan and launch one convolution.  The split-K buffer grows on demand only
// when may_alloc (not while a caller captures a graph); the preconditioner
// path reserves it when the network is attached.
int launch_conv(UserPrec &U, const ConvArgs &a, int mode, bool may_alloc)
{
  if(const int cc = conv_lds_cc(a)) {
    if(mode == M_POOL) return launch_conv_lds<M_POOL>(a, cc);
    if(mode == M_UPCAT) return launch_conv_lds<M_UPCAT>(a, cc);
    return launch_conv_lds<M_PLAIN>(a, cc);
  }
  const ConvPlan c = plan_conv((long)a.B * a.D * a.H * a.W, a.cout, a.K_pad);
  if((size_t)c.split_floats > U.split_size) {
    if(!may_alloc)
      return set_error(GDP_E_STATE, "convolution scratch not reserved (%ld floats)", c.split_floats);
    GDP_TRY(sync());
    dfree(U.split);
    U.split_size = 0;
    GDP_TRY(dalloc(&U.split, (size_t)c.split_floats));
    U.split_size = (size_t)c.split_floats;
  }
  if(mode == M_POOL) return launch_conv_mode<M_POOL>(a, c, U.split);
  if(mode == M_UPCAT) return launch_conv_mode<M_UPCAT>(a, c, U.split);
  return launch_conv_mode<M_PLAIN>(a, c, U.split);
}

// split-K scratch needed by a forward pass at batch B
size_t forward_split_floats(const UserPrec &U, int B)
{
  size_t m = 0;
  for(const auto &l : U.L) {
    long P = (long)B * U.vox(l.level);
    m = std::max(m, (size_t)plan_conv(P, l.cout, l.K_pad).split_floats);
  }
  return m;
}

// The convolutions, then (out_layer) the 1x1 output layer into U.z.
int forward(UserPrec &U, int B, const int *gate, int gmode, bool may_alloc, bool out_layer = true)
{
  for(size_t i = 0; i < U.L.size(); i++) {
    ConvArgs a = conv_args(U, U.L[i], B);
    a.out_bf = U.act[i + 1];
    a.gate = gate;
    a.gmode = gmode;
    GDP_TRY(launch_conv(U, a, U.L[i].mode, may_alloc));
  }
  long P = (long)B * U.vox(0);
  if(out_layer)
    hipLaunchKernelGGL(k_out_fwd, g1(P), dim3(256), 0, ctx().stream, P, 16, U.act.back(), U.ow, U.ob,
                       U.z, gate, gmode);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

// out[c] = sum over pixels of x[p][c] (times dz[p] when given); partials in U.part
template <typename T> int colsum(UserPrec &U, const T *x, long P, int C, const float *dz, float *out)
{
  const long n = P * C;
  const int nb = (int)((n + kColChunk - 1) / kColChunk);
  const size_t need = (size_t)nb * C;
  if(need > U.cpart_size) {
    dfree(U.cpart);
    GDP_TRY(dalloc(&U.cpart, need));
    U.cpart_size = need;
  }
  hipLaunchKernelGGL(k_colsum_part<T>, dim3(nb), dim3(256), 0, ctx().stream, n, C, x, dz, U.cpart);
  hipLaunchKernelGGL(k_colsum_fin, dim3(C), dim3(256), 0, ctx().stream, nb, C, U.cpart, out);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int backward(UserPrec &U, int B)
{
  hipStream_t s = ctx().stream;
  const long P0 = (long)B * U.vox(0);
  for(size_t t = 0; t < U.gact.size(); t++) {
    long n = (long)B * U.vox(U.act_lvl[t]) * U.act_ch[t];
    GDP_HIP(hipMemsetAsync(U.gact[t], 0, sizeof(float) * n, s));
  }
  // output 1x1: weight gradient sum_p dz[p] x[p][c], bias gradient sum_p dz[p]
  GDP_TRY(colsum(U, U.act.back(), P0, 16, U.dz, U.gow));
  GDP_TRY(colsum(U, U.dz, P0, 1, (const float *)nullptr, U.gob));
  hipLaunchKernelGGL(k_out_bwd_x, g1(P0 * 16), dim3(256), 0, s, P0, 16, U.dz, U.ow, U.gact.back());
  for(int i = (int)U.L.size() - 1; i >= 0; i--) {
    const ConvLayer &l = U.L[i];
    const int H = U.H >> l.level, W = U.W >> l.level, Dl = U.d3 ? U.D >> l.level : 1;
    const long P = (long)B * Dl * H * W;
    // rows of whole voxel octets: channel-major operands for the weight gradient
    const bool tr = W % 8 == 0;
    // G = relu'(y) * dY (BF16), and G^T
    if(tr)
      hipLaunchKernelGGL(k_relu_grad_t, dim3((unsigned)((P + kTrPix - 1) / kTrPix)), dim3(256), 0, s, P, l.cout,
                         U.gact[i + 1], U.act[i + 1], U.g, U.gt);
    else
      hipLaunchKernelGGL(k_relu_grad, g1(P * l.cout), dim3(256), 0, s, P * l.cout, U.gact[i + 1],
                         U.act[i + 1], U.g);
    GDP_TRY(colsum(U, U.g, P, l.cout, (const float *)nullptr, l.gb));
    // weight gradient: chunks of voxels small enough for ~2048 workgroups
    ConvArgs x = conv_args(U, l, B);
    const long tiles = (long)((l.cout + 15) / 16) * ((l.K_pad + 63) / 64);
    int chunk = 2048;
    while(chunk > 256 && (P + chunk - 1) / chunk * tiles < 2048) chunk >>= 1;
    int nchunks = (int)((P + chunk - 1) / chunk);
    size_t need = (size_t)nchunks * l.cout * l.K_pad;
    if(need > U.part_size) {
      dfree(U.part);
      GDP_TRY(dalloc(&U.part, need));
      U.part_size = need;
    }
    dim3 gw((unsigned)nchunks, (unsigned)((l.cout + 15) / 16), (unsigned)((l.K_pad + 63) / 64));
    const int wcc = tr ? wgrad_lds_cc(x) : 0;
    if(wcc) {
      // tiles per workgroup for ~1024 workgroups (fewer partial sums to add)
      const int TY = tile_ty(U.d3), TZ = tile_tz(U.d3);
      const long ntiles = (long)B * (W / kTileX) * (H / TY) * (U.d3 ? Dl / TZ : 1);
      const int nci = (l.cin_pad + wcc - 1) / wcc;
      const long other = (long)((l.cout + 15) / 16) * nci;
      const int ntile = (int)std::max(1L, std::min(16L, ntiles * other / 1024));
      const int nblk = (int)((ntiles + ntile - 1) / ntile);
      need = (size_t)nblk * l.cout * l.K_pad;
      if(need > U.part_size) {
        dfree(U.part);
        GDP_TRY(dalloc(&U.part, need));
        U.part_size = need;
      }
      nchunks = nblk;
      WgradLArgs wl{x, U.gt, U.part, wcc, ntile, (int)ntiles};
      const dim3 g((unsigned)nblk, (unsigned)((l.cout + 15) / 16), (unsigned)nci);
      const size_t lds = (size_t)(16 * 256 + wcc * (TY + 2) * (U.d3 ? TZ + 2 : 1) * kWgRow) * 2;
      if(wcc == 32) {
        if(l.mode == M_POOL) hipLaunchKernelGGL((k_wgrad_lds<M_POOL, 2>), g, dim3(256), lds, s, wl);
        else if(l.mode == M_UPCAT) hipLaunchKernelGGL((k_wgrad_lds<M_UPCAT, 2>), g, dim3(256), lds, s, wl);
        else hipLaunchKernelGGL((k_wgrad_lds<M_PLAIN, 2>), g, dim3(256), lds, s, wl);
      }
      else {
        if(l.mode == M_POOL) hipLaunchKernelGGL((k_wgrad_lds<M_POOL, 1>), g, dim3(256), lds, s, wl);
        else if(l.mode == M_UPCAT) hipLaunchKernelGGL((k_wgrad_lds<M_UPCAT, 1>), g, dim3(256), lds, s, wl);
        else hipLaunchKernelGGL((k_wgrad_lds<M_PLAIN, 1>), g, dim3(256), lds, s, wl);
      }
    }
    else if(tr) {
      u16 *xt = U.xt + kXtPad;
      const dim3 gt((unsigned)((P + kTrPix - 1) / kTrPix));
      if(l.mode == M_POOL) hipLaunchKernelGGL(k_input_t<M_POOL>, gt, dim3(256), 0, s, x, xt);
      else if(l.mode == M_UPCAT) hipLaunchKernelGGL(k_input_t<M_UPCAT>, gt, dim3(256), 0, s, x, xt);
      else hipLaunchKernelGGL(k_input_t<M_PLAIN>, gt, dim3(256), 0, s, x, xt);
      WgradTArgs wt{Dl, H, W, U.d3, l.cin_pad, l.cout, l.K_pad, P, U.gt, xt, U.part, chunk,
                    nchunks, (int)gw.y, (int)gw.z};
      const unsigned nb = (unsigned)(8 * ((nchunks + 7) / 8) * gw.y * gw.z);
      hipLaunchKernelGGL(k_wgrad_t, dim3(nb), dim3(256), 0, s, wt);
    }
    else {
      WgradArgs wa{};
      wa.x = x;
      wa.g = U.g;
      wa.chunk = chunk;
      wa.part = U.part;
      if(l.mode == M_POOL) hipLaunchKernelGGL(k_wgrad<M_POOL>, gw, dim3(256), 0, s, wa);
      else if(l.mode == M_UPCAT) hipLaunchKernelGGL(k_wgrad<M_UPCAT>, gw, dim3(256), 0, s, wa);
      else hipLaunchKernelGGL(k_wgrad<M_PLAIN>, gw, dim3(256), 0, s, wa);
    }
    long nw = (long)l.cout * l.K_pad;
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((nw + 31) / 32)), dim3(256), 0, s, nchunks, nw, U.part, l.gw);
    if(i == 0) continue;  // no gradient needed for the network input
    // input gradient: adjoint convolution (mirrored taps) on G
    ConvArgs d{};
    d.B = B;
    d.H = H;
    d.W = W;
    d.D = Dl;
    d.d3 = U.d3;
    d.cin_pad = l.cout;
    d.cout = l.cin_pad;
    d.K_pad = l.K2_pad;
    d.in = U.g;
    d.wt = l.wtT;
    d.out_f = U.dx;
    d.relu = 0;  // the taps are already mirrored in wtT
    GDP_TRY(launch_conv(U, d, M_PLAIN, true));
    long nx = P * l.cin_pad;
    if(l.mode == M_PLAIN)
      hipLaunchKernelGGL(k_add, g1(nx), dim3(256), 0, s, nx, U.dx, U.gact[l.in_tensor]);
    else if(l.mode == M_POOL)
      hipLaunchKernelGGL(k_pool_adj, g1(nx * (U.d3 ? 8 : 4)), dim3(256), 0, s, B, Dl, H, W, l.cin_pad, U.d3, U.dx,
                         U.gact[l.in_tensor]);
    else {
      long nt = (long)B * (U.d3 ? Dl / 2 : 1) * (H / 2) * (W / 2) * l.c1 + P * (l.cin_pad - l.c1);
      hipLaunchKernelGGL(k_upcat_adj, g1(nt), dim3(256), 0, s, B, Dl, H, W, U.d3, l.c1, l.cin_pad - l.c1,
                         U.dx, U.gact[l.in_tensor], U.gact[l.in2_tensor]);
    }
  }
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int prep_weights(UserPrec &U)
{
  for(auto &l : U.L) {
    long n = std::max((long)l.cout * l.K_pad, (long)l.cin_pad * l.K2_pad);
    hipLaunchKernelGGL(k_prep_weights, g1(n), dim3(256), 0, ctx().stream, l.cout, l.cin_pad, l.K_pad,
                       l.K2_pad, U.taps, l.w32, l.wt, l.wtT);
  }
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

}  // namespace

UserPrec::~UserPrec()
{
  for(auto &l : L) {
    zfree(l.w32);
    zfree(l.b32);
    zfree(l.gw);
    zfree(l.gb);
    zfree(l.mw);
    zfree(l.vw);
    zfree(l.mb);
    zfree(l.vb);
    zfree(l.wt);
    zfree(l.wtT);
  }
  for(auto p : act) dfree(p);
  for(auto p : gact) dfree(p);
  zfree(ow);
  zfree(ob);
  zfree(gow);
  zfree(gob);
  zfree(mow);
  zfree(vow);
  zfree(mob);
  zfree(vob);
  zfree(z);
  zfree(dz);
  zfree(cpart);
  zfree(target);
  zfree(mask);
  zfree(loss);
  zfree(dx);
  zfree(g);
  zfree(gt);
  zfree(xt);
  zfree(part);
  zfree(split);
  zfree(pix2dof);
  zfree(dof2pix);
  zfree(feat);
  zfree(r32);
  zfree(dscale);
  zfree(pc_part);
}

void user_prec_bind(UserPrec *u, const int *perm, const int *rperm, const double *rs)
{
  u->perm = perm;
  u->rperm = rperm;
  u->rs = rs;
}

// z = M^{-1} v through the network: map the solver vector to the grid
// (unpermute, unscale), normalise to unit RMS, forward pass, restore the
// scale and map back.
template <typename TV>
int user_prec_apply_t(UserPrec *u, const TV *in, TV *out, KCtl *C, int gmode)
{
  if(!u || !u->pix2dof) return set_error(GDP_E_STATE, "learned preconditioner not bound to a system");
  hipStream_t s = ctx().stream;
  const int *gate = (const int *)C;
  const int nparts = u->pc_nparts;
  const int npix = (int)u->vox(0);
  hipLaunchKernelGGL(k_pc_unscale<TV>, dim3(nparts), dim3(256), 0, s, npix, u->pix2dof, u->perm, u->rs,
                     u->dscale, in, u->r32, u->pc_part, gate, gmode);
  const int odd = u->odd && u->Bmax >= 2;
  hipLaunchKernelGGL(k_pc_gather, dim3((unsigned)std::min(1024, (npix + 255) / 256)), dim3(256), 0, s,
                     npix, u->n_dof, u->r32, u->pc_part, nparts, u->feat, u->act[0], odd, gate, gmode);
  // the output layer is evaluated in the scatter, at the DOF pixels only
  GDP_TRY(forward(*u, odd ? 2 : 1, gate, gmode, false, false));
  hipLaunchKernelGGL(k_pc_scatter<TV>, dim3((unsigned)std::min(1024, (u->n_dof + 255) / 256)), dim3(256), 0,
                     s, u->n_dof, u->n_dof, u->rperm, u->dof2pix, u->rs, u->dscale, u->act.back(), u->ow,
                     u->ob, u->pc_part, nparts, npix, odd, out, gate, gmode);
  GDP_HIP(hipGetLastError());
  return GDP_OK;
}

int user_prec_apply(UserPrec *u, const double *in, double *out, KCtl *C, int gmode)
{
  return user_prec_apply_t<double>(u, in, out, C, gmode);
}
int user_prec_apply(UserPrec *u, const float *in, float *out, KCtl *C, int gmode)
{
  return user_prec_apply_t<float>(u, in, out, C, gmode);
}

int user_prec_lsq(const UserPrec *u) { return u ? u->lsq : 0; }

UserPrec *solver_user_prec(gdp_solver *s);
void solver_set_user_prec(gdp_solver *s, UserPrec *u);

}  // namespace gdp

using namespace gdp;

// =============================================================================
// C-ABI
// =============================================================================
struct gdp_unet {
  std::unique_ptr<UserPrec> u;
};

static int unet_create(int D, int H, int W, int batch, unsigned seed, gdp_unet **out)
{
  GDP_TRY(conv_lds_attrs());
  if(!out || H <= 0 || W <= 0 || D <= 0 || H % 8 || W % 8 || (D > 1 && D % 8) || batch < 1)
    return set_error(GDP_E_ARG, "gdp_unet_create: grid sides must be positive multiples of 8");
  GDP_TRY(ensure_ready());
  auto *N = new gdp_unet;
  N->u.reset(new UserPrec);
  UserPrec &U = *N->u;
  U.H = H;
  U.W = W;
  U.D = D;
  U.d3 = D > 1;
  U.taps = U.d3 ? 27 : 9;
  const int taps = U.taps;
  U.Bmax = batch;
  // activations: 0 = input (8 channels)
  auto add_act = [&](int ch, int lvl) {
    U.act_ch.push_back(ch);
    U.act_lvl.push_back(lvl);
    return (int)U.act_ch.size() - 1;
  };
  add_act(8, 0);
  struct Spec {
    int mode, cout, lvl, in, in2;
  };
  // tensor indices: conv i writes tensor i+1
  const Spec spec[14] = {
    {M_PLAIN, 16, 0, 0, -1},   {M_PLAIN, 16, 0, 1, -1},                    // E0 -> s0 = t2
    {M_POOL, 32, 1, 2, -1},    {M_PLAIN, 32, 1, 3, -1},                    // E1 -> s1 = t4
    {M_POOL, 64, 2, 4, -1},    {M_PLAIN, 64, 2, 5, -1},                    // E2 -> s2 = t6
    {M_POOL, 128, 3, 6, -1},   {M_PLAIN, 128, 3, 7, -1},                   // E3 -> s3 = t8
    {M_UPCAT, 64, 2, 8, 6},    {M_PLAIN, 64, 2, 9, -1},                    // D2 -> t10
    {M_UPCAT, 32, 1, 10, 4},   {M_PLAIN, 32, 1, 11, -1},                   // D1 -> t12
    {M_UPCAT, 16, 0, 12, 2},   {M_PLAIN, 16, 0, 13, -1},                   // D0 -> t14
  };
  std::mt19937 rng(seed);
  for(int i = 0; i < 14; i++) {
    ConvLayer l;
    l.mode = spec[i].mode;
    l.cout = spec[i].cout;
    l.level = spec[i].lvl;
    l.in_tensor = spec[i].in;
    l.in2_tensor = spec[i].in2;
    if(l.mode == M_UPCAT) {
      l.c1 = U.act_ch[l.in_tensor];
      l.cin_pad = l.c1 + U.act_ch[l.in2_tensor];
      l.cin = l.cin_pad;
    }
    else {
      l.cin_pad = U.act_ch[l.in_tensor];
      l.cin = (i == 0) ? 2 : l.cin_pad;
    }
    l.K_pad = (int)rup((long)taps * l.cin_pad, 32);
    l.K2_pad = (int)rup((long)taps * l.cout, 32);
    add_act(l.cout, l.level);
    long nw = (long)l.cout * l.K_pad;
    std::vector<float> w(nw, 0.f), b(l.cout, 0.f);
    // He initialisation over the true fan-in
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / ((float)taps * l.cin)));
    for(int co = 0; co < l.cout; co++)
      for(int tap = 0; tap < taps; tap++)
        for(int ci = 0; ci < l.cin; ci++) w[(long)co * l.K_pad + tap * l.cin_pad + ci] = nd(rng);
    GDP_TRY(dalloc(&l.w32, nw));
    GDP_TRY(dalloc(&l.gw, nw));
    GDP_TRY(dalloc(&l.mw, nw));
    GDP_TRY(dalloc(&l.vw, nw));
    GDP_TRY(dalloc(&l.b32, l.cout));
    GDP_TRY(dalloc(&l.gb, l.cout));
    GDP_TRY(dalloc(&l.mb, l.cout));
    GDP_TRY(dalloc(&l.vb, l.cout));
    GDP_TRY(dalloc(&l.wt, nw));
    GDP_TRY(dalloc(&l.wtT, (size_t)l.cin_pad * l.K2_pad));
    GDP_TRY(h2d(l.w32, w.data(), nw));
    GDP_TRY(h2d(l.b32, b.data(), l.cout));
    GDP_HIP(hipMemsetAsync(l.mw, 0, sizeof(float) * nw, ctx().stream));
    GDP_HIP(hipMemsetAsync(l.vw, 0, sizeof(float) * nw, ctx().stream));
    GDP_HIP(hipMemsetAsync(l.mb, 0, sizeof(float) * l.cout, ctx().stream));
    GDP_HIP(hipMemsetAsync(l.vb, 0, sizeof(float) * l.cout, ctx().stream));
    U.L.push_back(l);
  }
  {
    std::vector<float> w(16), b(1, 0.f);
    std::normal_distribution<float> nd(0.f, std::sqrt(1.f / 16.f));
    for(auto &x : w) x = nd(rng);
    for(float **p : {&U.ow, &U.gow, &U.mow, &U.vow}) GDP_TRY(dalloc(p, 16));
    for(float **p : {&U.ob, &U.gob, &U.mob, &U.vob}) GDP_TRY(dalloc(p, 1));
    GDP_TRY(h2d(U.ow, w.data(), 16));
    GDP_TRY(h2d(U.ob, b.data(), 1));
    for(float *p : {U.mow, U.vow}) GDP_HIP(hipMemsetAsync(p, 0, 16 * sizeof(float), ctx().stream));
    for(float *p : {U.mob, U.vob}) GDP_HIP(hipMemsetAsync(p, 0, sizeof(float), ctx().stream));
  }
  size_t maxx = 0;
  for(size_t t = 0; t < U.act_ch.size(); t++) {
    size_t n = (size_t)batch * U.vox(U.act_lvl[t]) * U.act_ch[t];
    u16 *a = nullptr;
    float *g = nullptr;
    GDP_TRY(dalloc(&a, n));
    GDP_TRY(dalloc(&g, n));
    GDP_HIP(hipMemsetAsync(a, 0, n * sizeof(u16), ctx().stream));
    U.act.push_back(a);
    U.gact.push_back(g);
    maxx = std::max(maxx, n);
  }
  size_t maxin = 0;
  for(auto &l : U.L) maxin = std::max(maxin, (size_t)batch * U.vox(l.level) * l.cin_pad);
  GDP_TRY(dalloc(&U.dx, std::max(maxx, maxin)));
  GDP_TRY(dalloc(&U.g, maxx));
  // channel-major copies for the weight gradient (transpose tiles: cout <= 128, cin <= 192)
  for(auto &l : U.L)
    if(l.cout > 128 || l.cin_pad > 192) return set_error(GDP_E_UNSUPPORTED, "U-Net channel counts exceed the transpose tiles");
  GDP_TRY(dalloc(&U.gt, maxx));
  GDP_TRY(dalloc(&U.xt, maxin + 2 * kXtPad));
  GDP_HIP(hipMemsetAsync(U.xt, 0, (maxin + 2 * kXtPad) * sizeof(u16), ctx().stream));
  size_t P = (size_t)batch * U.vox(0);
  GDP_TRY(dalloc(&U.z, P));
  GDP_TRY(dalloc(&U.dz, P));
  GDP_TRY(dalloc(&U.target, P));
  GDP_TRY(dalloc(&U.mask, (size_t)U.vox(0)));
  GDP_TRY(dalloc(&U.loss, (size_t)batch));
  GDP_TRY(prep_weights(U));
  GDP_TRY(sync());
  *out = N;
  return GDP_OK;
}

extern "C" int gdp_unet_create(int H, int W, int batch, unsigned seed, gdp_unet **out)
{
  return unet_create(1, H, W, batch, seed, out);
}

extern "C" int gdp_unet_create_3d(int D, int H, int W, int batch, unsigned seed, gdp_unet **out)
{
  if(D < 8) return set_error(GDP_E_ARG, "gdp_unet_create_3d: depth must be a positive multiple of 8");
  return unet_create(D, H, W, batch, seed, out);
}

extern "C" int gdp_unet_destroy(gdp_unet *n)
{
  if(!n) return GDP_OK;
  if(ctx().ready) (void)sync();
  delete n;
  return GDP_OK;
}

extern "C" int gdp_unet_num_params(const gdp_unet *n, long *count)
{
  if(!n || !count) return set_error(GDP_E_ARG, "null argument");
  long c = 17;
  for(auto &l : n->u->L) c += (long)l.cout * n->u->taps * l.cin + l.cout;
  *count = c;
  return GDP_OK;
}

// Parameters as one FP32 array in layer order: for each convolution its
// weights [cout][9][cin] then biases [cout]; then the output weights [16] and
// bias [1].  get/set for checkpoints and for gradient checks.
static int params_io(gdp_unet *n, float *host, bool get, bool grads)
{
  UserPrec &U = *n->u;
  long off = 0;
  for(auto &l : U.L) {
    long nw = (long)l.cout * l.K_pad;
    std::vector<float> w(nw, 0.f), b(l.cout);
    float *dw = grads ? l.gw : l.w32, *db = grads ? l.gb : l.b32;
    GDP_TRY(d2h(w.data(), dw, nw));
    GDP_TRY(d2h(b.data(), db, l.cout));
    for(int co = 0; co < l.cout; co++)
      for(int tap = 0; tap < U.taps; tap++)
        for(int ci = 0; ci < l.cin; ci++) {
          float &slot = w[(long)co * l.K_pad + tap * l.cin_pad + ci];
          if(get) host[off] = slot;
          else slot = host[off];
          off++;
        }
    for(int co = 0; co < l.cout; co++) {
      if(get) host[off] = b[co];
      else b[co] = host[off];
      off++;
    }
    if(!get) {
      GDP_TRY(h2d(dw, w.data(), nw));
      GDP_TRY(h2d(db, b.data(), l.cout));
    }
  }
  float *ow = grads ? U.gow : U.ow, *ob = grads ? U.gob : U.ob;
  if(get) {
    GDP_TRY(d2h(host + off, ow, 16));
    GDP_TRY(d2h(host + off + 16, ob, 1));
  }
  else {
    GDP_TRY(h2d(ow, host + off, 16));
    GDP_TRY(h2d(ob, host + off + 16, 1));
    GDP_TRY(prep_weights(U));
  }
  return sync();
}

extern "C" int gdp_unet_get_params(gdp_unet *n, float *host)
{
  if(!n || !host) return set_error(GDP_E_ARG, "null argument");
  return params_io(n, host, true, false);
}

extern "C" int gdp_unet_set_params(gdp_unet *n, const float *host)
{
  if(!n || !host) return set_error(GDP_E_ARG, "null argument");
  return params_io(n, const_cast<float *>(host), false, false);
}

extern "C" int gdp_unet_get_grads(gdp_unet *n, float *host)
{
  if(!n || !host) return set_error(GDP_E_ARG, "null argument");
  return params_io(n, host, true, true);
}

// Forward pass on a batch of grids: r, feat (per pixel) -> z (FP32).
extern "C" int gdp_unet_forward(gdp_unet *n, int batch, const float *r, const float *feat,
                                float *z_out)
{
  if(!n || !r || !feat || batch < 1 || batch > n->u->Bmax) return set_error(GDP_E_ARG, "gdp_unet_forward");
  UserPrec &U = *n->u;
  const long HW = U.vox(0), P = batch * HW;
  float *dr = nullptr, *df = nullptr;
  GDP_TRY(dalloc(&dr, P));
  GDP_TRY(dalloc(&df, P));
  int rc = h2d(dr, r, P);
  if(!rc) rc = h2d(df, feat, P);
  if(!rc) {
    hipLaunchKernelGGL(k_make_input, g1(P), dim3(256), 0, ctx().stream, P, dr, df, U.act[0]);
    rc = forward(U, batch, nullptr, 0, true);
  }
  if(!rc && z_out) rc = d2h(z_out, U.z, P);
  dfree(dr);
  dfree(df);
  return rc;
}

// One training step on a batch (scale-invariant loss on the masked grid).
// Returns the mean loss in *loss.  lr = 0 computes gradients only.
extern "C" int gdp_unet_train_step(gdp_unet *n, int batch, const float *r, const float *feat,
                                   const float *target, const float *mask, float lr,
                                   float *loss)
{
  if(!n || !r || !feat || !target || !mask || batch < 1 || batch > n->u->Bmax)
    return set_error(GDP_E_ARG, "gdp_unet_train_step");
  UserPrec &U = *n->u;
  hipStream_t s = ctx().stream;
  const long HW = U.vox(0), P = batch * HW;
  float *dr = nullptr, *df = nullptr;
  GDP_TRY(dalloc(&dr, P));
  GDP_TRY(dalloc(&df, P));
  int rc = h2d(dr, r, P);
  if(!rc) rc = h2d(df, feat, P);
  if(!rc) rc = h2d(U.target, target, P);
  if(!rc) rc = h2d(U.mask, mask, HW);
  if(!rc) {
    hipLaunchKernelGGL(k_make_input, g1(P), dim3(256), 0, s, P, dr, df, U.act[0]);
    rc = forward(U, batch, nullptr, 0, true);
  }
  if(!rc) {
    hipLaunchKernelGGL(k_loss, dim3(batch), dim3(256), 0, s, (int)HW, U.z, U.target, U.mask, U.dz,
                       U.loss);
    rc = backward(U, batch);
  }
  if(!rc && lr > 0.f) {
    U.adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - std::pow(b1, (float)U.adam_t), bc2 = 1.f - std::pow(b2, (float)U.adam_t);
    for(auto &l : U.L) {
      long nw = (long)l.cout * l.K_pad;
      hipLaunchKernelGGL(k_adam, g1(nw), dim3(256), 0, s, nw, l.w32, l.gw, l.mw, l.vw, lr, b1, b2, eps, bc1, bc2, 0.f);
      hipLaunchKernelGGL(k_adam, g1(l.cout), dim3(256), 0, s, (long)l.cout, l.b32, l.gb, l.mb, l.vb, lr, b1, b2, eps, bc1, bc2, 0.f);
    }
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, s, 16L, U.ow, U.gow, U.mow, U.vow, lr, b1, b2, eps, bc1, bc2, 0.f);
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, s, 1L, U.ob, U.gob, U.mob, U.vob, lr, b1, b2, eps, bc1, bc2, 0.f);
    rc = prep_weights(U);
  }
  if(!rc && loss) {
    std::vector<float> h(batch);
    rc = d2h(h.data(), U.loss, batch);
    float m = 0.f;
    for(float v : h) m += v;
    *loss = m / batch;
  }
  dfree(dr);
  dfree(df);
  return rc;
}

// Bind the network as the FGMRES preconditioner of a solver: pix2dof maps grid
// pixels to (renumbered) DOF indices (-1: not a DOF), feat holds the per-pixel
// coefficient feature.  Used when the solver's Algorithm is FGMRES (9).
extern "C" int gdp_unet_attach(gdp_unet *n, gdp_solver *s, int n_dof, const int *pix2dof,
                               const float *feat, const double *dscale)
{
  if(!n || !s || !pix2dof || !feat || n_dof < 1) return set_error(GDP_E_ARG, "gdp_unet_attach");
  UserPrec &U = *n->u;
  const long npix = U.vox(0);
  std::vector<int> d2p(n_dof, -1);
  for(long p = 0; p < npix; p++) {
    int d = pix2dof[p];
    if(d >= n_dof) return set_error(GDP_E_ARG, "pix2dof[%ld] = %d out of range", p, d);
    if(d >= 0) d2p[d] = (int)p;
  }
  for(int d = 0; d < n_dof; d++)
    if(d2p[d] < 0) return set_error(GDP_E_ARG, "DOF %d has no grid pixel", d);
  dfree(U.pix2dof);
  dfree(U.dof2pix);
  dfree(U.feat);
  dfree(U.r32);
  dfree(U.dscale);
  dfree(U.pc_part);
  GDP_TRY(dalloc(&U.pix2dof, npix));
  GDP_TRY(dalloc(&U.dof2pix, n_dof));
  GDP_TRY(dalloc(&U.feat, npix));
  GDP_TRY(dalloc(&U.r32, npix));
  U.pc_nparts = (int)std::min<long>(256, (npix + 255) / 256);
  GDP_TRY(dalloc(&U.pc_part, U.pc_nparts));
  std::vector<u16> fb(npix);
  for(long p = 0; p < npix; p++) fb[p] = f2bf(feat[p]);
  GDP_TRY(h2d(U.pix2dof, pix2dof, npix));
  GDP_TRY(h2d(U.dof2pix, d2p.data(), n_dof));
  GDP_TRY(h2d(U.feat, fb.data(), npix));
  if(dscale) {
    for(int d = 0; d < n_dof; d++)
      if(!(dscale[d] > 0.)) return set_error(GDP_E_ARG, "dscale[%d] must be positive", d);
    GDP_TRY(dalloc(&U.dscale, n_dof));
    GDP_TRY(h2d(U.dscale, dscale, n_dof));
  }
  U.n_dof = n_dof;
  // reserve the convolution scratch of the preconditioner's passes
  const size_t need = std::max(forward_split_floats(U, 1), forward_split_floats(U, std::min(2, U.Bmax)));
  if(need > U.split_size) {
    GDP_TRY(sync());
    dfree(U.split);
    U.split_size = 0;
    GDP_TRY(dalloc(&U.split, need));
    U.split_size = need;
  }
  solver_set_user_prec(s, &U);
  return sync();
}

extern "C" int gdp_unet_detach(gdp_solver *s)
{
  if(!s) return set_error(GDP_E_ARG, "null solver");
  solver_set_user_prec(s, nullptr);
  return GDP_OK;
}

// Copy activation tensor t (0 = network input, i + 1 = output of convolution
// i, 15 = network output z as one FP32 channel) of the last forward pass to
// the host as FP32, NHWC with the stored (padded) channel count.
extern "C" int gdp_unet_get_activation(gdp_unet *n, int t, int batch, float *host, long *count)
{
  if(!n || batch < 1 || batch > n->u->Bmax) return set_error(GDP_E_ARG, "gdp_unet_get_activation");
  UserPrec &U = *n->u;
  const int nt = (int)U.act.size();
  if(t < 0 || t > nt) return set_error(GDP_E_ARG, "activation index %d out of range", t);
  if(t == nt) {
    long c = (long)batch * U.vox(0);
    if(count) *count = c;
    if(host) GDP_TRY(d2h(host, U.z, (size_t)c));
    return GDP_OK;
  }
  long c = (long)batch * U.vox(U.act_lvl[t]) * U.act_ch[t];
  if(count) *count = c;
  if(!host) return GDP_OK;
  std::vector<u16> h(c);
  GDP_TRY(d2h(h.data(), U.act[t], (size_t)c));
  for(long i = 0; i < c; i++) {
    unsigned int u = ((unsigned int)h[i]) << 16;
    std::memcpy(host + i, &u, 4);
  }
  return GDP_OK;
}

// Preconditioner options: bit 0 = apply the odd part (N(r) - N(-r))/2 of the
// network (one batch-2 pass; the network must have been created with
// batch >= 2); bit 1 = in the two-stage preconditioner, scale the network's
// output z by alpha = <v, Az> / <Az, Az> before the ILU(0) correction.
extern "C" int gdp_unet_set_mode(gdp_unet *n, int flags)
{
  if(!n) return set_error(GDP_E_ARG, "null network");
  if((flags & 1) && n->u->Bmax < 2) return set_error(GDP_E_ARG, "odd mode needs a network created with batch >= 2");
  n->u->odd = flags & 1;
  n->u->lsq = (flags >> 1) & 1;
  return GDP_OK;
}
