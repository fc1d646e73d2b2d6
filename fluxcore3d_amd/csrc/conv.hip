// conv.hip — 3x3x3 convolutions of the learned preconditioner on MFMA (gfx950).
//
// Implicit GEMM: out[p, co] = sum_k W[co, k] X[p + o(tap), ci], k = tap * Cin + ci,
// M = Cout (16-row tiles), N = 16 consecutive positions along W, K = 27 Cin
// (padded to a multiple of 32 with zero weights). One wavefront computes a
// 16-position tile for all Cout with v_mfma_f32_16x16x32_bf16 (BF16 operands,
// FP32 accumulation). Activations are channels-last BF16 [B, D, H, W, C] with
// C a multiple of 8, so one lane's 8 K-values are one 16-byte load.
//
// A two-source input (decoder stages) concatenates a half-resolution tensor
// upsampled by nearest neighbour (channels [0, c0)) with a full-resolution
// skip tensor (channels [c0, c0 + c1)) on the fly.
//
// The data gradient is the same kernel with flipped, transposed weights; the
// weight gradient is a GEMM over positions (K = positions) per (Cout tile,
// tap, Cin tile), with per-chunk partials folded in a fixed order.
#include "fc_common.hpp"

#include <algorithm>

using namespace fc;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct ConvShape {
  int B, D, H, W;  // output (and full-resolution input) extents
  int c0, c1;      // channels of source 0 and source 1
  int up0;         // source 0 is half resolution (nearest upsample)
  int cout, kpad;
};

__device__ __forceinline__ int tap_dz(int t) { return t / 9 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return (t / 3) % 3 - 1; }
__device__ __forceinline__ int tap_dx(int t) { return t % 3 - 1; }

// 8 consecutive input channels [ci0, ci0+8) at full-resolution position
// (b, d, h, w); zeros outside the volume or beyond the 27 taps.
__device__ __forceinline__ bf16x8 load_b8(const __bf16* __restrict__ x0, const __bf16* __restrict__ x1,
                                          const ConvShape& s, int b, int d, int h, int w, int ci0) {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
  if (d < 0 || d >= s.D || h < 0 || h >= s.H || w < 0 || w >= s.W) return z;
  if (ci0 < s.c0) {
    if (s.up0) {
      const int D2 = s.D >> 1, H2 = s.H >> 1, W2 = s.W >> 1;
      const int64_t p = (((int64_t)b * D2 + (d >> 1)) * H2 + (h >> 1)) * W2 + (w >> 1);
      return *(const bf16x8*)(x0 + p * s.c0 + ci0);
    }
    const int64_t p = (((int64_t)b * s.D + d) * s.H + h) * s.W + w;
    return *(const bf16x8*)(x0 + p * s.c0 + ci0);
  }
  const int64_t p = (((int64_t)b * s.D + d) * s.H + h) * s.W + w;
  return *(const bf16x8*)(x1 + p * s.c1 + (ci0 - s.c0));
}

// One wavefront computes NT consecutive 16-position tiles along W for all
// output channels (MT 16-row tiles): per 32-wide K step it loads NT input
// fragments and MT weight fragments and issues MT x NT MFMAs.
// act: 0 linear, 1 leaky ReLU (slope)
template <int MT, int NT>
__global__ __launch_bounds__(256) void conv_fwd_kernel(const __bf16* __restrict__ x0, const __bf16* __restrict__ x1,
                                                       const __bf16* __restrict__ wt, ConvShape s, int act,
                                                       float slope, __bf16* __restrict__ y16,
                                                       float* __restrict__ y32) {
  const int lane = threadIdx.x & 63;
  const int64_t tiles_w = (s.W + 16 * NT - 1) / (16 * NT);
  const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t n_tiles = (int64_t)s.B * s.D * s.H * tiles_w;
  if (tile >= n_tiles) return;
  const int wt_i = (int)(tile % tiles_w);
  int64_t r = tile / tiles_w;
  const int h = (int)(r % s.H); r /= s.H;
  const int d = (int)(r % s.D);
  const int b = (int)(r / s.D);
  const int w0 = wt_i * 16 * NT;
  const int col = lane & 15, grp = lane >> 4;
  const int cin = s.c0 + s.c1;
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int ksteps = s.kpad / 32;
  for (int ks = 0; ks < ksteps; ++ks) {
    const int k0 = ks * 32 + grp * 8;
    const int tap = k0 / cin, ci0 = k0 - tap * cin;
    bf16x8 bv[NT];
    if (tap < 27) {
#pragma unroll
      for (int n = 0; n < NT; ++n)
        bv[n] = load_b8(x0, x1, s, b, d + tap_dz(tap), h + tap_dy(tap), w0 + 16 * n + col + tap_dx(tap), ci0);
    } else {
#pragma unroll
      for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[n][j] = (__bf16)0.0f;
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      // rows past cout read the last row (valid memory); their results are dropped
      const int co = min(m * 16 + col, s.cout - 1);
      const bf16x8 av = *(const bf16x8*)(wt + (int64_t)co * s.kpad + k0);
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[n], acc[m][n], 0, 0, 0);
    }
  }
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const int w = w0 + 16 * n + col;
    if (w >= s.W) continue;
    const int64_t p = (((int64_t)b * s.D + d) * s.H + h) * s.W + w;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int co = m * 16 + grp * 4;
      if (co >= s.cout) continue;
      float v[4];
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        float a = acc[m][n][rr];
        if (act == 1 && a < 0.f) a *= slope;
        v[rr] = a;
      }
      if (y32) {
        *(f32x4*)(y32 + p * s.cout + co) = (f32x4){v[0], v[1], v[2], v[3]};
      }
      if (y16) {
        bf16x4 o;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) o[rr] = (__bf16)v[rr];
        *(bf16x4*)(y16 + p * s.cout + co) = o;
      }
    }
  }
}

// Same convolution with the input staged in LDS. A block computes up to kLdsDZ
// output planes of 4 x 64 positions (one h row of four 16-position tiles per
// wavefront), walking along d with a ring of three input planes of 6 x 66
// voxels: each output plane loads one new input plane (16-byte pieces of 8
// channels, zero outside the volume) and every tap reads its B fragments from
// LDS. The LDS offsets of each lane group's K steps are tabulated once per
// block. The cached kernel above re-gathers each input voxel for all 27 taps of
// every position.
constexpr int kLdsTH = 4, kLdsTW = 64, kLdsPH = kLdsTH + 2, kLdsPW = kLdsTW + 2, kLdsDZ = 8;  // max planes per block
constexpr int kLdsPlane = kLdsPH * kLdsPW;  // voxels of one input plane
constexpr int kLdsMaxK = 128;               // K-step table entries (kpad / 8 <= 27 * 3 + 3)
constexpr int kLoadBatch = 8;               // 16-byte input loads a thread has in flight while staging

// tap = kk / nch for kk < 256 and nch <= 8 by a multiply-shift with the host's
// magic = ceil(2^16 / nch) (checked on the host for every kk used)
__device__ __forceinline__ int div_small(int kk, int magic) { return (kk * magic) >> 16; }

template <int MT>
__global__ __launch_bounds__(256) void conv_fwd_lds_kernel(const __bf16* __restrict__ x0,
                                                           const __bf16* __restrict__ x1,
                                                           const __bf16* __restrict__ wt, ConvShape s, int act,
                                                           float slope, int magic, int dz_block,
                                                           __bf16* __restrict__ y16, float* __restrict__ y32) {
  extern __shared__ bf16x8 ring[];  // [plane slot 3][chunk of 8 channels][hy 6][wx 66]
  __shared__ int2 ktab[kLdsMaxK];   // per 8-wide K group: {tile offset without the plane, dz}
  constexpr int NT = kLdsTW / 16;
  const int nch = (s.c0 + s.c1) >> 3, nch0 = s.c0 >> 3;
  const int slab = nch * kLdsPlane;  // one plane slot
  const int tiles_w = (s.W + kLdsTW - 1) / kLdsTW, tiles_h = (s.H + kLdsTH - 1) / kLdsTH;
  const int tiles_d = (s.D + dz_block - 1) / dz_block;
  unsigned bid = blockIdx.x;
  const int tw = (int)(bid % (unsigned)tiles_w); bid /= (unsigned)tiles_w;
  const int th = (int)(bid % (unsigned)tiles_h); bid /= (unsigned)tiles_h;
  const int td = (int)(bid % (unsigned)tiles_d);
  const int b = (int)(bid / (unsigned)tiles_d);
  const int w0 = tw * kLdsTW, h0 = th * kLdsTH, d0 = td * dz_block, d1 = min(d0 + dz_block, s.D);
  const int kgroups = s.kpad >> 3;
  for (int kk = threadIdx.x; kk < kgroups; kk += 256) {
    // K past 27 taps carries zero weights: any finite tile values serve, so the
    // padding reads tap 26 (tile values are finite activations)
    const int tap0 = div_small(kk, magic), chunk = kk - tap0 * nch, tap = min(tap0, 26);
    const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
    ktab[kk] = make_int2((chunk * kLdsPH + ty) * kLdsPW + tx, tz);
  }
  // item `it` of input plane dd (absolute): 8 channels of one voxel, zero outside the volume
  auto fetch = [&](int dd, int it) -> bf16x8 {
    const int wx = it % kLdsPW, r1 = it / kLdsPW;
    const int hy = r1 % kLdsPH, ch = r1 / kLdsPH;
    const int hh = h0 + hy - 1, ww = w0 + wx - 1;
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
    if ((unsigned)dd < (unsigned)s.D && (unsigned)hh < (unsigned)s.H && (unsigned)ww < (unsigned)s.W) {
      if (ch < nch0) {
        const unsigned pos = s.up0 ? ((unsigned)(b * (s.D >> 1) + (dd >> 1)) * (unsigned)(s.H >> 1) + (unsigned)(hh >> 1)) *
                                             (unsigned)(s.W >> 1) + (unsigned)(ww >> 1)
                                   : ((unsigned)(b * s.D + dd) * (unsigned)s.H + (unsigned)hh) * (unsigned)s.W + (unsigned)ww;
        v = *(const bf16x8*)(x0 + (pos * (unsigned)s.c0 + (unsigned)(ch << 3)));
      } else {
        const unsigned pos = ((unsigned)(b * s.D + dd) * (unsigned)s.H + (unsigned)hh) * (unsigned)s.W + (unsigned)ww;
        v = *(const bf16x8*)(x1 + (pos * (unsigned)s.c1 + (unsigned)((ch - nch0) << 3)));
      }
    }
    return v;
  };
  // input plane dd (absolute) into ring slot dd mod 3; a thread's loads are all issued
  // before its first LDS store (one memory latency per batch of kLoadBatch, not per item)
  auto load_plane = [&](int dd) {
    bf16x8* dst = ring + ((dd + 3) % 3) * slab;
    const int n_items = nch * kLdsPlane;
    for (int it0 = threadIdx.x; it0 < n_items; it0 += 256 * kLoadBatch) {
      bf16x8 v[kLoadBatch];
#pragma unroll
      for (int u = 0; u < kLoadBatch; ++u)
        if (it0 + 256 * u < n_items) v[u] = fetch(dd, it0 + 256 * u);
#pragma unroll
      for (int u = 0; u < kLoadBatch; ++u)
        if (it0 + 256 * u < n_items) dst[it0 + 256 * u] = v[u];
    }
  };
  load_plane(d0 - 1);
  load_plane(d0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int h = h0 + wv;
  const int col = lane & 15, grp = lane >> 4;
  const int ksteps = s.kpad >> 5;
  const __bf16* wrow[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m)  // rows past cout read the last row (valid memory); results dropped
    wrow[m] = wt + (unsigned)min(m * 16 + col, s.cout - 1) * (unsigned)s.kpad + (unsigned)(grp << 3);
  const int lane_off = wv * kLdsPW + col;
  for (int d = d0; d < d1; ++d) {
    load_plane(d + 1);
    __syncthreads();  // planes d-1, d, d+1 and the K table are in LDS
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int s0 = (d + 2) % 3;  // ring slot of input plane d - 1
    bf16x8 av_next[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) av_next[m] = *(const bf16x8*)wrow[m];
    for (int ks = 0; ks < ksteps; ++ks) {
      bf16x8 av[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) av[m] = av_next[m];
      if (ks + 1 < ksteps) {  // next step's weights in flight during this step's MFMAs
#pragma unroll
        for (int m = 0; m < MT; ++m) av_next[m] = *(const bf16x8*)(wrow[m] + ((ks + 1) << 5));
      }
      const int2 e = ktab[(ks << 2) + grp];
      int sl = s0 + e.y;
      sl -= sl >= 3 ? 3 : 0;
      const bf16x8* row = ring + sl * slab + e.x + lane_off;
      bf16x8 bv[NT];
#pragma unroll
      for (int n = 0; n < NT; ++n) bv[n] = row[16 * n];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[n], acc[m][n], 0, 0, 0);
    }
    if (h < s.H) {
      const unsigned prow = ((unsigned)(b * s.D + d) * (unsigned)s.H + (unsigned)h) * (unsigned)s.W;
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int w = w0 + 16 * n + col;
        if (w >= s.W) continue;
        const unsigned p = prow + (unsigned)w;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int co = m * 16 + grp * 4;
          if (co >= s.cout) continue;
          float v[4];
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            float a = acc[m][n][rr];
            if (act == 1 && a < 0.f) a *= slope;
            v[rr] = a;
          }
          if (y32) *(f32x4*)(y32 + (p * (unsigned)s.cout + (unsigned)co)) = (f32x4){v[0], v[1], v[2], v[3]};
          if (y16) {
            bf16x4 o;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) o[rr] = (__bf16)v[rr];
            *(bf16x4*)(y16 + (p * (unsigned)s.cout + (unsigned)co)) = o;
          }
        }
      }
    }
    __syncthreads();  // every wavefront is done with plane d - 1 before it is overwritten
  }
}

// The LDS-staged convolution for R output planes per wavefront (inputs of up to 16
// channels, up to 16 output channels). With one output plane per wavefront every B
// fragment read from LDS feeds MT MFMAs (one for 16 output channels), and the layer is
// bound by LDS reads. Here a wavefront computes planes d .. d+R-1 of its row: the K
// steps run over (ty, tx, channel chunk) for each tz, and the B fragment of input
// plane d-1+e serves every output plane r with tz = e - r in [0, 2]: R+2 fragment reads
// feed 3R MFMAs (R = 4: half the LDS reads per MFMA). The block's ring holds the R+2
// input planes; each step of R output planes loads R new ones.
template <int MT, int R>
__global__ __launch_bounds__(256) void conv_fwd_lds_r_kernel(const __bf16* __restrict__ x0,
                                                             const __bf16* __restrict__ wt, ConvShape s, int act,
                                                             float slope, int dz_block, __bf16* __restrict__ y16,
                                                             float* __restrict__ y32) {
  extern __shared__ bf16x8 ring[];  // [plane slot R+2][chunk of 8 channels][hy 6][wx 66]
  __shared__ int2 ktab[kLdsMaxK];   // per 8-wide group of a tz block: {tile offset, weight offset or -1}
  constexpr int NT = kLdsTW / 16, NS = R + 2;
  const int nch = s.c0 >> 3, cin = s.c0;
  const int slab = nch * kLdsPlane;
  const int tiles_w = (s.W + kLdsTW - 1) / kLdsTW, tiles_h = (s.H + kLdsTH - 1) / kLdsTH;
  const int tiles_d = (s.D + dz_block - 1) / dz_block;
  unsigned bid = blockIdx.x;
  const int tw = (int)(bid % (unsigned)tiles_w); bid /= (unsigned)tiles_w;
  const int th = (int)(bid % (unsigned)tiles_h); bid /= (unsigned)tiles_h;
  const int td = (int)(bid % (unsigned)tiles_d);
  const int b = (int)(bid / (unsigned)tiles_d);
  const int w0 = tw * kLdsTW, h0 = th * kLdsTH, d0 = td * dz_block, d1 = min(d0 + dz_block, s.D);
  const int ngroups = 9 * nch;            // 8-wide K groups of one tz block
  const int jsteps = (ngroups + 3) >> 2;  // 32-wide K steps per tz block
  for (int kg = threadIdx.x; kg < 4 * jsteps; kg += 256) {
    int2 e = make_int2(0, -1);  // past the block's groups: zero weights, any finite tile value
    if (kg < ngroups) {
      const int t9 = kg / nch, ch = kg - t9 * nch;
      e = make_int2((ch * kLdsPH + t9 / 3) * kLdsPW + t9 % 3, t9 * cin + (ch << 3));
    }
    ktab[kg] = e;
  }
  // input planes [da, db] (absolute) into ring slots (dd - d0 + 1) mod NS; a thread's
  // loads are all issued before its first LDS store (one memory latency per batch of
  // kLoadBatch items, not per item)
  auto load_planes = [&](int da, int db) {
    const int per = nch * kLdsPlane, n_items = (db - da + 1) * per;
    for (int it0 = threadIdx.x; it0 < n_items; it0 += 256 * kLoadBatch) {
      bf16x8 v[kLoadBatch];
#pragma unroll
      for (int u = 0; u < kLoadBatch; ++u) {
        const int it = it0 + 256 * u;
        if (it >= n_items) continue;
        const int pl = it / per, rem = it - pl * per, dd = da + pl;
        const int wx = rem % kLdsPW, r1 = rem / kLdsPW;
        const int hy = r1 % kLdsPH, ch = r1 / kLdsPH;
        const int hh = h0 + hy - 1, ww = w0 + wx - 1;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[u][j] = (__bf16)0.0f;
        if ((unsigned)dd < (unsigned)s.D && (unsigned)hh < (unsigned)s.H && (unsigned)ww < (unsigned)s.W) {
          const unsigned pos = ((unsigned)(b * s.D + dd) * (unsigned)s.H + (unsigned)hh) * (unsigned)s.W + (unsigned)ww;
          v[u] = *(const bf16x8*)(x0 + (pos * (unsigned)cin + (unsigned)(ch << 3)));
        }
      }
#pragma unroll
      for (int u = 0; u < kLoadBatch; ++u) {
        const int it = it0 + 256 * u;
        if (it >= n_items) continue;
        const int pl = it / per;
        ring[((da + pl - d0 + 1) % NS) * slab + (it - pl * per)] = v[u];
      }
    }
  };
  load_planes(d0 - 1, min(d0 + R, d1));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int h = h0 + wv;
  const int col = lane & 15, grp = lane >> 4;
  const __bf16* wrow[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m)  // rows past cout read the last row (valid memory); results dropped
    wrow[m] = wt + (unsigned)min(m * 16 + col, s.cout - 1) * (unsigned)s.kpad;
  const int lane_off = wv * kLdsPW + col;
  const int tzk = 9 * cin;  // weight offset between tz blocks
  bf16x8 zero8;
#pragma unroll
  for (int j = 0; j < 8; ++j) zero8[j] = (__bf16)0.0f;
  for (int d = d0; d < d1; d += R) {
    if (d > d0 && d + 1 <= d1) load_planes(d + 1, min(d + R, d1));  // d-1 and d are in the ring already
    __syncthreads();  // planes d-1 .. d+R and the K table are in LDS
    f32x4 acc[R][MT][NT];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[r][m][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < jsteps; ++j) {
      const int2 e = ktab[(j << 2) + grp];
      bf16x8 av[3][MT];
#pragma unroll
      for (int tz = 0; tz < 3; ++tz)
#pragma unroll
        for (int m = 0; m < MT; ++m) av[tz][m] = e.y >= 0 ? *(const bf16x8*)(wrow[m] + tz * tzk + e.y) : zero8;
#pragma unroll
      for (int pe = 0; pe < NS; ++pe) {
        const bf16x8* row = ring + ((d - d0 + pe) % NS) * slab + e.x + lane_off;  // input plane d-1+pe
        bf16x8 bv[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) bv[n] = row[16 * n];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int tz = pe - r;
          if (tz < 0 || tz > 2) continue;
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n)
              acc[r][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[tz][m], bv[n], acc[r][m][n], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (d + r >= d1 || h >= s.H) continue;
      const unsigned prow = ((unsigned)(b * s.D + d + r) * (unsigned)s.H + (unsigned)h) * (unsigned)s.W;
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int w = w0 + 16 * n + col;
        if (w >= s.W) continue;
        const unsigned p = prow + (unsigned)w;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int co = m * 16 + grp * 4;
          if (co >= s.cout) continue;
          float v[4];
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            float a = acc[r][m][n][rr];
            if (act == 1 && a < 0.f) a *= slope;
            v[rr] = a;
          }
          if (y32) *(f32x4*)(y32 + (p * (unsigned)s.cout + (unsigned)co)) = (f32x4){v[0], v[1], v[2], v[3]};
          if (y16) {
            bf16x4 o;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) o[rr] = (__bf16)v[rr];
            *(bf16x4*)(y16 + (p * (unsigned)s.cout + (unsigned)co)) = o;
          }
        }
      }
    }
    __syncthreads();  // every wavefront is done with planes d-1 .. d+R-2 before they are overwritten
  }
}

// W'[ci][t'][co] = W[co][26 - t'][ci] (data-gradient weights), zero-padded rows.
__global__ void flip_weights_kernel(const __bf16* __restrict__ w, int cout, int cin, int kpad,
                                    __bf16* __restrict__ wf, int kpad_f) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t)cin * kpad_f;
  if (i >= n) return;
  const int ci = (int)(i / kpad_f), k = (int)(i % kpad_f);
  __bf16 v = (__bf16)0.0f;
  if (k < 27 * cout) {
    const int t2 = k / cout, co = k % cout;
    v = w[(int64_t)co * kpad + (26 - t2) * cin + ci];
  }
  wf[i] = v;
}

// dW partials: one wave per (co tile, ci tile, tap, chunk of rows); K = positions
// along W in groups of 32 (8 per lane), lane-gathered BF16 loads.
__global__ __launch_bounds__(64) void conv_dw_kernel(const __bf16* __restrict__ x0, const __bf16* __restrict__ x1,
                                                     const __bf16* __restrict__ dy, ConvShape s, int rows_per_chunk,
                                                     int n_chunks, float* __restrict__ partial) {
  const int lane = threadIdx.x;
  const int cin = s.c0 + s.c1;
  const int n_mt = (s.cout + 15) / 16, n_nt = cin / 16 + ((cin % 16) ? 1 : 0);
  int64_t id = blockIdx.x;
  const int chunk = (int)(id % n_chunks); id /= n_chunks;
  const int tap = (int)(id % 27); id /= 27;
  const int nt = (int)(id % n_nt);
  const int mt = (int)(id / n_nt);
  const int col = lane & 15, grp = lane >> 4;
  const int co = mt * 16 + col;
  const int ci = nt * 16 + col;
  const int dz = tap_dz(tap), dyy = tap_dy(tap), dx = tap_dx(tap);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int64_t n_rows = (int64_t)s.B * s.D * s.H;
  const int64_t r_begin = (int64_t)chunk * rows_per_chunk;
  const int64_t r_end = min(n_rows, r_begin + rows_per_chunk);
  for (int64_t row = r_begin; row < r_end; ++row) {
    const int h = (int)(row % s.H);
    const int d = (int)((row / s.H) % s.D);
    const int b = (int)(row / ((int64_t)s.H * s.D));
    for (int w0 = 0; w0 < s.W; w0 += 32) {
      bf16x8 av, bv;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int w = w0 + grp * 8 + j;
        __bf16 a = (__bf16)0.0f, bb = (__bf16)0.0f;
        if (w < s.W) {
          const int64_t p = (((int64_t)b * s.D + d) * s.H + h) * s.W + w;
          if (co < s.cout) a = dy[p * s.cout + co];
          const int dd = d + dz, hh = h + dyy, ww = w + dx;
          if (ci < cin && dd >= 0 && dd < s.D && hh >= 0 && hh < s.H && ww >= 0 && ww < s.W) {
            if (ci < s.c0) {
              if (s.up0) {
                const int D2 = s.D >> 1, H2 = s.H >> 1, W2 = s.W >> 1;
                const int64_t q = (((int64_t)b * D2 + (dd >> 1)) * H2 + (hh >> 1)) * W2 + (ww >> 1);
                bb = x0[q * s.c0 + ci];
              } else {
                const int64_t q = (((int64_t)b * s.D + dd) * s.H + hh) * s.W + ww;
                bb = x0[q * s.c0 + ci];
              }
            } else {
              const int64_t q = (((int64_t)b * s.D + dd) * s.H + hh) * s.W + ww;
              bb = x1[q * s.c1 + (ci - s.c0)];
            }
          }
        }
        av[j] = a;
        bv[j] = bb;
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
  }
  // acc[r] = dW[co = mt*16 + grp*4 + r][tap][ci = nt*16 + col]
  const int64_t base = (int64_t)chunk * s.cout * 27 * cin;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int c_o = mt * 16 + grp * 4 + rr;
    if (c_o < s.cout && ci < cin) partial[base + ((int64_t)c_o * 27 + tap) * cin + ci] = acc[rr];
  }
}

__global__ void fold_chunks_kernel(const float* __restrict__ partial, int n_chunks, int64_t n, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < n_chunks; ++c) s += partial[(int64_t)c * n + i];
  out[i] = s;
}

}  // namespace

// Dynamic LDS the staged kernel may take: the device's per-block limit less the static
// K table, at most 80 KB (two blocks per CU on a 160 KB CU; a third would not fit).
// Cached per device; 60 KB where the attribute is unavailable.
static size_t lds_cap() {
  static size_t cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 61440;
  if (!cache[dev]) {
    int v = 0;
    size_t cap = 61440;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 4096)
      cap = std::min<size_t>((size_t)v - sizeof(int2) * kLdsMaxK, 81920);
    cache[dev] = cap;
  }
  return cache[dev];
}

extern "C" {

int fc_conv_kpad(int cin) { return (27 * cin + 31) / 32 * 32; }

int fc_conv3d_fwd(const void* x0, int c0, int up0, const void* x1, int c1, const void* w, int cout, int B, int D,
                  int H, int W, int act, float slope, void* y_bf16, float* y_f32, hipStream_t stream) {
  FC_REQUIRE(x0 && w, "conv3d_fwd: null input or weights");
  FC_REQUIRE(c0 > 0 && c0 % 8 == 0 && c1 >= 0 && c1 % 8 == 0, "conv3d_fwd: channel counts must be multiples of 8");
  FC_REQUIRE(c1 == 0 || x1, "conv3d_fwd: second source missing");
  FC_REQUIRE(cout > 0 && cout % 4 == 0 && cout <= 64, "conv3d_fwd: cout must be a multiple of 4 and <= 64");
  FC_REQUIRE(B > 0 && D > 0 && H > 0 && W > 0, "conv3d_fwd: extents must be positive");
  FC_REQUIRE(!up0 || (D % 2 == 0 && H % 2 == 0 && W % 2 == 0), "conv3d_fwd: upsampled source needs even extents");
  FC_REQUIRE(y_bf16 || y_f32, "conv3d_fwd: no output");
  ConvShape s{B, D, H, W, c0, c1, up0, cout, fc_conv_kpad(c0 + c1)};
  const int mt = (cout + 15) / 16;
  // LDS-staged kernel where rows are long enough and the three input planes fit in
  // lds_cap(): up to 4 channel chunks (76 KB with the K table, 2 blocks per CU on
  // gfx950's 160 KB) where the device allows it
  const size_t lds = (size_t)((c0 + c1) / 8) * 3 * kLdsPlane * sizeof(bf16x8);
  const int64_t elems = (int64_t)B * D * H * W * (c0 + c1 > cout ? c0 + c1 : cout);
  // R = 4 output planes per wavefront: one source of up to 16 channels, up to 16 outputs
  constexpr int kR = 4;
  const size_t lds_r = (size_t)(c0 / 8) * (kR + 2) * kLdsPlane * sizeof(bf16x8);
  if (c1 == 0 && !up0 && c0 <= 16 && cout <= 16 && W >= kLdsTW && lds_r <= lds_cap() && elems < ((int64_t)1 << 31)) {
    // output planes per block: a multiple of R up to 8 while keeping >= 2048 blocks
    const int64_t per_plane = (int64_t)B * ((H + kLdsTH - 1) / kLdsTH) * ((W + kLdsTW - 1) / kLdsTW);
    const int dz_block = per_plane * D >= 2048 * 8 ? 8 : kR;
    const int64_t blocks = per_plane * ((D + dz_block - 1) / dz_block);
    hipLaunchKernelGGL((conv_fwd_lds_r_kernel<1, kR>), dim3((unsigned)blocks), dim3(256), lds_r, stream,
                       (const __bf16*)x0, (const __bf16*)w, s, act, slope, dz_block, (__bf16*)y_bf16, y_f32);
    return check_launch("conv3d_fwd/lds_r");
  }
  if (W >= kLdsTW && lds <= lds_cap() && elems < ((int64_t)1 << 31) && fc_conv_kpad(c0 + c1) / 8 <= kLdsMaxK) {
    // output planes per block: up to kLdsDZ (fewer halo reloads) while keeping >= 2048 blocks
    const int64_t per_plane = (int64_t)B * ((H + kLdsTH - 1) / kLdsTH) * ((W + kLdsTW - 1) / kLdsTW);
    const int dz_block = (int)std::max<int64_t>(1, std::min<int64_t>(kLdsDZ, per_plane * D / 2048));
    const int64_t blocks = per_plane * ((D + dz_block - 1) / dz_block);
    const int nch = (c0 + c1) / 8, magic = (65536 + nch - 1) / nch;
    // kk / nch by multiply-shift: exact for every kk the kernel forms (kk < kpad / 8 <= 27 nch)
    for (int kk = 0; kk < fc_conv_kpad(c0 + c1) / 8; ++kk)
      FC_REQUIRE((kk * magic) >> 16 == kk / nch, "conv3d_fwd: reciprocal of %d channel chunks inexact", nch);
#define FC_CONV_LDS(M)                                                                                          \
  hipLaunchKernelGGL((conv_fwd_lds_kernel<M>), dim3((unsigned)blocks), dim3(256), lds, stream, (const __bf16*)x0, \
                     (const __bf16*)x1, (const __bf16*)w, s, act, slope, magic, dz_block, (__bf16*)y_bf16, y_f32)
    switch (mt) {
      case 1: FC_CONV_LDS(1); break;
      case 2: FC_CONV_LDS(2); break;
      case 3: FC_CONV_LDS(3); break;
      default: FC_CONV_LDS(4); break;
    }
#undef FC_CONV_LDS
    return check_launch("conv3d_fwd/lds");
  }
  const int nt = W >= 64 ? 4 : (W >= 32 ? 2 : 1);
  const int64_t tiles = (int64_t)B * D * H * ((W + 16 * nt - 1) / (16 * nt));
  const unsigned grid = (unsigned)((tiles + 3) / 4);
#define FC_CONV_LAUNCH(M, N)                                                                              \
  hipLaunchKernelGGL((conv_fwd_kernel<M, N>), dim3(grid), dim3(256), 0, stream, (const __bf16*)x0,      \
                     (const __bf16*)x1, (const __bf16*)w, s, act, slope, (__bf16*)y_bf16, y_f32)
#define FC_CONV_NT(M)                          \
  switch (nt) {                                \
    case 1: FC_CONV_LAUNCH(M, 1); break;       \
    case 2: FC_CONV_LAUNCH(M, 2); break;       \
    default: FC_CONV_LAUNCH(M, 4); break;      \
  }
  switch (mt) {
    case 1: FC_CONV_NT(1); break;
    case 2: FC_CONV_NT(2); break;
    case 3: FC_CONV_NT(3); break;
    default: FC_CONV_NT(4); break;
  }
#undef FC_CONV_NT
#undef FC_CONV_LAUNCH
  return check_launch("conv3d_fwd");
}

int fc_conv3d_flip_weights(const void* w, int cout, int cin, void* w_flip, hipStream_t stream) {
  FC_REQUIRE(w && w_flip && cout > 0 && cin > 0, "conv3d_flip_weights: bad arguments");
  const int kpad = fc_conv_kpad(cin), kpad_f = fc_conv_kpad(cout);
  const int64_t n = (int64_t)cin * kpad_f;
  hipLaunchKernelGGL(flip_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     (const __bf16*)w, cout, cin, kpad, (__bf16*)w_flip, kpad_f);
  return check_launch("conv3d_flip_weights");
}

size_t fc_conv3d_dw_workspace_bytes(int cout, int cin, int n_chunks) {
  return sizeof(float) * (size_t)n_chunks * cout * 27 * cin;
}

// dW[co][tap][ci] (FP32, [cout][27][cin]) for the forward call with the same sources.
int fc_conv3d_bwd_weight(const void* x0, int c0, int up0, const void* x1, int c1, const void* dy, int cout, int B,
                         int D, int H, int W, int n_chunks, float* dw, void* workspace, size_t ws_bytes,
                         hipStream_t stream) {
  FC_REQUIRE(x0 && dy && dw, "conv3d_bwd_weight: null pointer argument");
  FC_REQUIRE(c0 > 0 && c0 % 8 == 0 && c1 >= 0 && c1 % 8 == 0 && (c1 == 0 || x1), "conv3d_bwd_weight: bad channels");
  FC_REQUIRE(n_chunks > 0, "conv3d_bwd_weight: n_chunks must be > 0");
  const int cin = c0 + c1;
  const size_t need = fc_conv3d_dw_workspace_bytes(cout, cin, n_chunks);
  FC_REQUIRE(workspace && ws_bytes >= need, "conv3d_bwd_weight: workspace of %zu bytes required", need);
  ConvShape s{B, D, H, W, c0, c1, up0, cout, fc_conv_kpad(cin)};
  const int64_t n_rows = (int64_t)B * D * H;
  const int rows_per_chunk = (int)((n_rows + n_chunks - 1) / n_chunks);
  const int n_mt = (cout + 15) / 16, n_nt = (cin + 15) / 16;
  const unsigned grid = (unsigned)((int64_t)n_mt * n_nt * 27 * n_chunks);
  hipLaunchKernelGGL(conv_dw_kernel, dim3(grid), dim3(64), 0, stream, (const __bf16*)x0, (const __bf16*)x1,
                     (const __bf16*)dy, s, rows_per_chunk, n_chunks, (float*)workspace);
  if (int rc = check_launch("conv3d_bwd_weight")) return rc;
  const int64_t n = (int64_t)cout * 27 * cin;
  hipLaunchKernelGGL(fold_chunks_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     (const float*)workspace, n_chunks, n, dw);
  return check_launch("conv3d_bwd_weight/fold");
}

}  // extern "C"
