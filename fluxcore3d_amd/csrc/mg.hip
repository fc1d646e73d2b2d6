// mg.hip — geometric multigrid for the implicit thermal operator A = I + dt L.
//
// L is the 7-point finite-volume operator of nlevel_mg_cuda.py:418-507
// (harmonic-mean diffusion + first-order upwind), with the boundary ghosts of
// mg_kernels.cu:37-44: the high face of every axis mirrors to n-2, the low
// face mirrors to 1 except on the flow axis with the inlet at index 0, where
// the ghost is the node itself.
//
// Entry points:
//  * fc_mg_mop_fused / fc_mg_jacobi_smooth — the two reference kernels
//    (mg_kernels.cu:55-124), same arguments;
//  * fc_mg_build_coeffs — the coefficient fields of one level, every float
//    operation rounded on its own (same values as the eager tensor ops);
//  * fc_mg_restrict_fields — 2x2x2 means for the hierarchy build;
//  * fc_mg_sweep — fused weighted-Jacobi sweep x_out = x + D^-1 (b - A x);
//  * fc_mg_residual_restrict — coarse rhs = mean_8((b - A x) * mask);
//  * fc_mg_prolong_correct — x = (x + 0.75 * trilinear(x_c)) * mask;
//  * fc_mg_coarse_solve — all coarsest-level sweeps in one workgroup;
//  * FP64 vector kernels of the flexible CG outer loop, with wavefront DPP
//    reductions and device-resident scalars.
#include "fc_common.hpp"

using namespace fc;

namespace {

constexpr int kBlock = 256;
// FP64 reductions run a capped grid (grid-stride loops): one partial per block,
// folded by a single workgroup; the fixed grid keeps the summation order fixed.
// (16384 and 32768 made the 256^3 FCG solve 0.5 % and 2 % slower: the fold grows.)
constexpr unsigned kMaxPartials = 8192;

struct Level {
  int nx, ny, nz;
  int ax, out_idx;
};

struct Coef {
  const float *aE, *aW, *aN, *aS, *aU, *aD, *diagL, *invD;
  // face form (fc_common.hpp face_hi/face_lo); the kernels templated on FACE
  // read these three arrays instead of the six per-cell ones and diagL
  const float *fx, *fy, *fz;
  // optional [nx][ny] row table: 0 where every cell of the z-row (i, j) has zero
  // coefficients and a zero right-hand side, so that its iterate, residual and
  // search directions stay exactly 0; the stencil and vector kernels skip such rows
  const uint8_t* rows;
};

// false for a skipped row (its outputs keep the zeros they were allocated with)
// The skip table holds one entry per run of 64 consecutive cells (flat index >> 6).
constexpr int kSkipShift = 6;
__device__ __forceinline__ bool cell_live(const Coef& c, int idx) {
  return c.rows == nullptr || c.rows[(unsigned)idx >> kSkipShift] != 0;
}

__device__ __forceinline__ int lo_ghost(const Level& g, int this_ax) {
  return (g.ax != this_ax || g.out_idx == 0) ? 1 : 0;
}

// Cell counts stay below 2^31 (make_level checks it), so index math is 32-bit:
// a 64-bit division per cell costs more than the rest of a stencil update.
__device__ __forceinline__ void unflat(const Level& g, int idx, int& i, int& j, int& k) {
  const unsigned nynz = (unsigned)g.ny * (unsigned)g.nz, u = (unsigned)idx;
  i = (int)(u / nynz);
  const unsigned rem = u - (unsigned)i * nynz;
  j = (int)(rem / (unsigned)g.nz);
  k = (int)(rem - (unsigned)j * (unsigned)g.nz);
}

// Offsets of the six stencil neighbours of (i,j,k) with the reference's ghosts
// (mg_kernels.cu:37-44): high faces mirror to n-2, low faces to index 1 (or to the
// node itself on the flow axis with the inlet at 0).
struct Nbr {
  int e, w, n, s, u, d;
};
__device__ __forceinline__ Nbr neighbours(const Level& g, int i, int j, int k) {
  const int nynz = g.ny * g.nz;
  Nbr o;
  o.e = (i < g.nx - 1 ? 1 : -1) * nynz;
  o.w = (i > 0 ? -1 : lo_ghost(g, 0)) * nynz;
  o.n = (j < g.ny - 1 ? 1 : -1) * g.nz;
  o.s = (j > 0 ? -1 : lo_ghost(g, 1)) * g.nz;
  o.u = k < g.nz - 1 ? 1 : -1;
  o.d = k > 0 ? -1 : lo_ghost(g, 2);
  return o;
}

// The six coefficients and diagL of cell (i,j,k). In face form diagL is summed
// in build_coeffs_kernel's order, so both forms give the same bits wherever
// fc_mg_build_faces found the face form exact.
struct C7 {
  float e, w, n, s, u, d, diag;
};
template <bool FACE>
__device__ __forceinline__ C7 coef_at(const Coef& c, const Level& g, int idx, int i, int j, int k) {
  C7 a;
  if constexpr (FACE) {
    const int sx = g.ny * g.nz;
    a.e = c.fx[face_hi(idx, i, g.nx, sx)];
    a.w = c.fx[face_lo(idx, i, g.nx, sx, lo_ghost(g, 0))];
    a.n = c.fy[face_hi(idx, j, g.ny, g.nz)];
    a.s = c.fy[face_lo(idx, j, g.ny, g.nz, lo_ghost(g, 1))];
    a.u = c.fz[face_hi(idx, k, g.nz, 1)];
    a.d = c.fz[face_lo(idx, k, g.nz, 1, lo_ghost(g, 2))];
    a.diag = ((((a.e + a.w) + a.n) + a.s) + a.u) + a.d;
  } else {
    a.e = c.aE[idx]; a.w = c.aW[idx]; a.n = c.aN[idx]; a.s = c.aS[idx]; a.u = c.aU[idx]; a.d = c.aD[idx];
    a.diag = c.diagL[idx];
  }
  return a;
}

// Lo = T + dt L(T) from the seven values, with the reference's operand order. Every
// kernel forms Lo, the sweep update and the prolongation through these helpers,
// compiled without multiply-add contraction: each product and sum is rounded on
// its own, so fused and separate kernels give the same bits whatever the compiler
// would fuse in each context.
__device__ __forceinline__ float lo_of(const C7& a, float dt, float TC, float TE, float TW, float TN, float TS,
                                       float TU, float TD) {
#pragma clang fp contract(off)
  const float LT = a.diag * TC - a.e * TE - a.w * TW - a.n * TN - a.s * TS - a.u * TU - a.d * TD;
  return TC + dt * LT;
}

// x + invD (b - Lo)
__device__ __forceinline__ float sweep_update(float x, float invD, float b, float lo) {
#pragma clang fp contract(off)
  return x + invD * (b - lo);
}

template <bool FACE, typename TT>
__device__ __forceinline__ float stencil_lo(const TT* __restrict__ T, const Coef& c, const Level& g, float dt, int idx,
                                            int i, int j, int k) {
  const Nbr o = neighbours(g, i, j, k);
  const C7 a = coef_at<FACE>(c, g, idx, i, j, k);
  return lo_of(a, dt, (float)T[idx], (float)T[idx + o.e], (float)T[idx + o.w], (float)T[idx + o.n],
               (float)T[idx + o.s], (float)T[idx + o.u], (float)T[idx + o.d]);
}

// Streaming accesses: values read or written once per kernel bypass the cache's
// retention, leaving L2 to the iterate and coefficient lines that neighbouring
// cells read again.
__device__ __forceinline__ float ld_once(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_once(float* p, float v) { __builtin_nontemporal_store(v, p); }

// First sweep from x = 0 (x1 = invD b), written the way sweep_kernel writes it.
__device__ __forceinline__ float zero_sweep(const Coef& c, const float* __restrict__ b, int idx) {
  return sweep_update(0.f, c.invD[idx], b[idx], 0.f);
}

__global__ __launch_bounds__(kBlock) void mop_kernel(const float* __restrict__ T, Coef c, Level g, float dt,
                                                     float* __restrict__ Lo) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= g.nx * g.ny * g.nz) return;
  int i, j, k;
  unflat(g, idx, i, j, k);
  Lo[idx] = stencil_lo<false>(T, c, g, dt, idx, i, j, k);
}

__global__ __launch_bounds__(kBlock) void jacobi_kernel(float* __restrict__ x, const float* __restrict__ b,
                                                        const float* __restrict__ Lo,
                                                        const float* __restrict__ invD, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n) return;
  x[idx] = sweep_update(x[idx], invD[idx], b[idx], Lo[idx]);
}

// x_out = x + invD (b - A x): the two reference kernels in one pass.
template <bool FACE>
__global__ __launch_bounds__(kBlock) void sweep_kernel(const float* __restrict__ x, float* __restrict__ x_out,
                                                       const float* __restrict__ b, Coef c, Level g, float dt,
                                                       int zero_guess) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= g.nx * g.ny * g.nz) return;
  int i, j, k;
  if (!cell_live(c, idx)) return;
  unflat(g, idx, i, j, k);
  if (zero_guess) {  // x = 0: A x = 0
    x_out[idx] = zero_sweep(c, b, idx);
    return;
  }
  const float lo = stencil_lo<FACE>(x, c, g, dt, idx, i, j, k);
  st_once(x_out + idx, sweep_update(x[idx], ld_once(c.invD + idx), ld_once(b + idx), lo));
}

// The first two sweeps from x = 0 in one pass: x1 = invD b is formed at the seven
// stencil points from b and invD instead of being written and read back;
// x_out = x1 + invD (b - A x1), the same bits as two sweep_kernel launches.
template <bool FACE>
__global__ __launch_bounds__(kBlock) void sweep2_from_zero_kernel(float* __restrict__ x_out,
                                                                  const float* __restrict__ b, Coef c, Level g,
                                                                  float dt) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= g.nx * g.ny * g.nz) return;
  int i, j, k;
  if (!cell_live(c, idx)) return;
  unflat(g, idx, i, j, k);
  const Nbr o = neighbours(g, i, j, k);
  const C7 a = coef_at<FACE>(c, g, idx, i, j, k);
  const float xc = zero_sweep(c, b, idx);
  const float lo = lo_of(a, dt, xc, zero_sweep(c, b, idx + o.e), zero_sweep(c, b, idx + o.w),
                         zero_sweep(c, b, idx + o.n), zero_sweep(c, b, idx + o.s), zero_sweep(c, b, idx + o.u),
                         zero_sweep(c, b, idx + o.d));
  st_once(x_out + idx, sweep_update(xc, c.invD[idx], b[idx], lo));
}

// coarse rhs: mean of the 8 fine residuals (b - A x) * mask (odd extents truncated).
// A block covers 32 coarse cells of one coarse row: its four wavefronts take the
// four fine rows (2I + di, 2J + dj), one fine cell per lane along k (contiguous
// loads); the residuals meet in LDS and lane K of the first wavefront adds the 8
// of coarse cell K in the order d = 4 di + 2 dj + dk.
template <bool FACE>
__global__ __launch_bounds__(kBlock) void residual_restrict_kernel(const float* __restrict__ x,
                                                                   const float* __restrict__ b,
                                                                   const float* __restrict__ mask, Coef c, Level g,
                                                                   float dt, float* __restrict__ bc, int cx, int cy,
                                                                   int cz, int kchunks) {
  static_assert(kBlock == 256, "four fine rows of 64 cells per block");
  __shared__ float res[4][64];
  const int row = blockIdx.x / kchunks;  // coarse (I, J)
  const int K0 = (blockIdx.x - row * kchunks) * 32;
  const int I = row / cy, J = row - I * cy;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = 2 * I + (wave >> 1), j = 2 * J + (wave & 1), k = 2 * K0 + lane;
  const int idx = (i * g.ny + j) * g.nz + k;
  // the whole block leaves when the 64-cell runs holding its fine cells are all skipped
  // (each lane tests its own cell's run, in parallel): the right-hand side of its coarse
  // cells is then 0, which the buffer already holds
  if (c.rows != nullptr && !__syncthreads_or(k < 2 * cz && c.rows[(unsigned)idx >> kSkipShift] != 0)) return;
  float r = 0.f;
  if (k < 2 * cz) {
    const float lo = stencil_lo<FACE>(x, c, g, dt, idx, i, j, k);
    r = (ld_once(b + idx) - lo) * ld_once(mask + idx);
  }
  res[wave][lane] = r;
  __syncthreads();
  const int K = K0 + threadIdx.x;
  if (threadIdx.x < 32 && K < cz) {
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 8; ++d) s += res[d >> 1][2 * threadIdx.x + (d & 1)];
    bc[(I * cy + J) * cz + K] = s / 8.f;
  }
}

// linear interpolation source (align_corners = False) along one axis; scale =
// n_in / n_out, divided once on the host (the same IEEE quotient)
__device__ __forceinline__ void lin_src(int o, float scale, int n_in, int& i0, int& i1, float& l0, float& l1) {
#pragma clang fp contract(off)
  float src = scale * ((float)o + 0.5f) - 0.5f;
  src = src < 0.f ? 0.f : src;
  i0 = min((int)src, n_in - 1);
  i1 = min(i0 + 1, n_in - 1);
  l1 = src - (float)i0;
  l0 = 1.f - l1;
}

// Coarse correction at fine (i,j,k): trilinear interpolation of the coarse
// iterate (F.interpolate, align_corners=False) through X(I, J, K).
template <typename F>
__device__ __forceinline__ float prolong_at(const F& X, int i, int j, int k, int cx, int cy, int cz, float sx,
                                            float sy, float sz) {
#pragma clang fp contract(off)
  int i0, i1, j0, j1, k0, k1;
  float a0, a1, b0, b1, c0, c1;
  lin_src(i, sx, cx, i0, i1, a0, a1);
  lin_src(j, sy, cy, j0, j1, b0, b1);
  lin_src(k, sz, cz, k0, k1, c0, c1);
  return a0 * (b0 * (c0 * X(i0, j0, k0) + c1 * X(i0, j0, k1)) + b1 * (c0 * X(i0, j1, k0) + c1 * X(i0, j1, k1))) +
         a1 * (b0 * (c0 * X(i1, j0, k0) + c1 * X(i1, j0, k1)) + b1 * (c0 * X(i1, j1, k0) + c1 * X(i1, j1, k1)));
}

// x = (x + alpha e) * mask (nlevel_mg_cuda.py:300-303)
__device__ __forceinline__ float corrected(float x, float e, float alpha, float m) {
#pragma clang fp contract(off)
  return (x + alpha * e) * m;
}

__global__ __launch_bounds__(kBlock) void prolong_correct_kernel(const float* __restrict__ xc, float* __restrict__ x,
                                                                 const float* __restrict__ mask, int nx, int ny,
                                                                 int nz, int cx, int cy, int cz, float sx, float sy,
                                                                 float sz, float alpha,
                                                                 const uint8_t* __restrict__ rows) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= nx * ny * nz) return;
  const unsigned nynz = (unsigned)ny * (unsigned)nz, u = (unsigned)idx;
  const int i = (int)(u / nynz);
  const unsigned rem = u - (unsigned)i * nynz;
  const int j = (int)(rem / (unsigned)nz), k = (int)(rem - (unsigned)j * (unsigned)nz);
  if (rows && !rows[u >> kSkipShift]) return;
  auto X = [&](int a, int b2, int c2) { return xc[(a * cy + b2) * cz + c2]; };
  x[idx] = corrected(x[idx], prolong_at(X, i, j, k, cx, cy, cz, sx, sy, sz), alpha, mask[idx]);
}

// Two fine cells (i, j, k), (i, j, k + 1) per thread, k even: 8-byte loads and
// stores, and the coarse values the two interpolations share are loaded once.
__global__ __launch_bounds__(kBlock) void prolong_correct2_kernel(const float* __restrict__ xc, float* __restrict__ x,
                                                                  const float* __restrict__ mask, int nx, int ny,
                                                                  int nz, int cx, int cy, int cz, float sx, float sy,
                                                                  float sz, float alpha,
                                                                  const uint8_t* __restrict__ rows) {
  const int hz = nz >> 1;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nx * ny * hz) return;
  const unsigned nyhz = (unsigned)ny * (unsigned)hz, u = (unsigned)t;
  const int i = (int)(u / nyhz);
  const unsigned rem = u - (unsigned)i * nyhz;
  const int j = (int)(rem / (unsigned)hz), k = 2 * (int)(rem - (unsigned)j * (unsigned)hz);
  const int idx = (i * ny + j) * nz + k;  // even (nz is), so idx and idx + 1 share a 64-cell run
  if (rows && !rows[(unsigned)idx >> kSkipShift]) return;
  auto X = [&](int a, int b2, int c2) { return xc[(a * cy + b2) * cz + c2]; };
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 xv = __builtin_nontemporal_load(reinterpret_cast<const f2*>(x + idx));
  const f2 mv = __builtin_nontemporal_load(reinterpret_cast<const f2*>(mask + idx));
  f2 out;
  out.x = corrected(xv.x, prolong_at(X, i, j, k, cx, cy, cz, sx, sy, sz), alpha, mv.x);
  out.y = corrected(xv.y, prolong_at(X, i, j, k + 1, cx, cy, cz, sx, sy, sz), alpha, mv.y);
  __builtin_nontemporal_store(out, reinterpret_cast<f2*>(x + idx));
}

// All sweeps of the coarsest level in one workgroup (global ping-pong, the
// level is small enough to stay in L2); x holds the result on return.
template <bool FACE>
__global__ __launch_bounds__(1024) void coarse_solve_kernel(float* __restrict__ x, float* __restrict__ tmp,
                                                            const float* __restrict__ b, Coef c, Level g, float dt,
                                                            int sweeps, int zero_guess) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  float* src = x;
  float* dst = tmp;
  for (int s = 0; s < sweeps; ++s) {
    for (int idx = threadIdx.x; idx < (int)n; idx += blockDim.x) {
      if (s == 0 && zero_guess) {
        dst[idx] = zero_sweep(c, b, idx);
      } else {
        int i, j, k;
        unflat(g, idx, i, j, k);
        const float lo = stencil_lo<FACE>(src, c, g, dt, idx, i, j, k);
        dst[idx] = sweep_update(src[idx], c.invD[idx], b[idx], lo);
      }
    }
    __syncthreads();
    float* t = src; src = dst; dst = t;
  }
  if (src != x) {
    for (int64_t idx = threadIdx.x; idx < n; idx += blockDim.x) x[idx] = src[idx];
  }
}

// The same sweeps for a coarsest level of at most kCoarseLdsCells cells: the
// iterate ping-pongs in LDS and each thread keeps the coefficients, invD, b and
// neighbour offsets of its cells in registers, so a sweep costs LDS reads and a
// barrier instead of L2 round trips. Same arithmetic, same bits.
constexpr int kCoarseThreads = 1024, kCoarseCpt = 4, kCoarseLdsCells = kCoarseThreads * kCoarseCpt;

template <bool FACE>
__global__ __launch_bounds__(kCoarseThreads) void coarse_solve_lds_kernel(float* __restrict__ x,
                                                                          const float* __restrict__ b, Coef c,
                                                                          Level g, float dt, int sweeps,
                                                                          int zero_guess) {
  __shared__ float S[2][kCoarseLdsCells];
  const int n = g.nx * g.ny * g.nz;
  C7 co[kCoarseCpt];
  Nbr nb[kCoarseCpt];
  float inv[kCoarseCpt], bb[kCoarseCpt];
#pragma unroll
  for (int u = 0; u < kCoarseCpt; ++u) {
    const int idx = (int)threadIdx.x + u * kCoarseThreads;
    if (idx < n) {
      int i, j, k;
      unflat(g, idx, i, j, k);
      nb[u] = neighbours(g, i, j, k);
      co[u] = coef_at<FACE>(c, g, idx, i, j, k);
      inv[u] = c.invD[idx];
      bb[u] = b[idx];
      S[0][idx] = zero_guess ? 0.f : x[idx];
    }
  }
  __syncthreads();
  int cur = 0;
  for (int s = 0; s < sweeps; ++s) {
    const float* src = S[cur];
    float* dst = S[cur ^ 1];
#pragma unroll
    for (int u = 0; u < kCoarseCpt; ++u) {
      const int idx = (int)threadIdx.x + u * kCoarseThreads;
      if (idx < n) {
        if (s == 0 && zero_guess) {
          dst[idx] = sweep_update(0.f, inv[u], bb[u], 0.f);
        } else {
          const Nbr& o = nb[u];
          const float lo = lo_of(co[u], dt, src[idx], src[idx + o.e], src[idx + o.w], src[idx + o.n],
                                 src[idx + o.s], src[idx + o.u], src[idx + o.d]);
          dst[idx] = sweep_update(src[idx], inv[u], bb[u], lo);
        }
      }
    }
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int u = 0; u < kCoarseCpt; ++u) {
    const int idx = (int)threadIdx.x + u * kCoarseThreads;
    if (idx < n) x[idx] = S[cur][idx];
  }
}

// ---- coefficient fields (nlevel_mg_cuda.py:418-507), contraction off -------
#pragma clang fp contract(off)
__device__ __forceinline__ float nb_field(const float* A, const Level& g, int i, int j, int k, int dim, int step) {
  int a[3] = {i, j, k};
  const int n[3] = {g.nx, g.ny, g.nz};
  if (step > 0) a[dim] = a[dim] < n[dim] - 1 ? a[dim] + 1 : a[dim] - 1;
  else a[dim] = a[dim] > 0 ? a[dim] - 1 : lo_ghost(g, dim);
  return A[((int64_t)a[0] * g.ny + a[1]) * g.nz + a[2]];
}

__global__ __launch_bounds__(kBlock) void build_coeffs_kernel(const float* __restrict__ kf, const float* __restrict__ rcp,
                                                              const float* __restrict__ u, const float* __restrict__ v,
                                                              const float* __restrict__ w, Level g, float cd, float ca,
                                                              float* aE, float* aW, float* aN, float* aS, float* aU,
                                                              float* aD, float* diagL) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= (int64_t)g.nx * g.ny * g.nz) return;
  int i, j, k;
  unflat(g, (int)idx, i, j, k);
  const float kc = kf[idx];
  const float inv_rcp = 1.0f / (rcp[idx] + 1e-30f);
  float out[6];
  const float vel[3] = {u[idx], v[idx], w[idx]};
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const int dim = f >> 1, step = (f & 1) ? -1 : +1;  // E W N S U D
    const float kn = nb_field(kf, g, i, j, k, dim, step);
    const float num = (2.0f * kc) * kn;
    const float den = (kc + kn) + 1e-30f;
    const float harm = num / den;
    const float diff = (cd * harm) * inv_rcp;
    // upwind: W/S/D take max(u,0), E/N/U take max(-u,0)
    const float up = step > 0 ? fmaxf(-vel[dim], 0.0f) : fmaxf(vel[dim], 0.0f);
    out[f] = diff + ca * up;
  }
  aE[idx] = out[0]; aW[idx] = out[1]; aN[idx] = out[2]; aS[idx] = out[3]; aU[idx] = out[4]; aD[idx] = out[5];
  diagL[idx] = ((((out[0] + out[1]) + out[2]) + out[3]) + out[4]) + out[5];
}

// inv_diagM = omega * (1 / max(1 + dt diagL, 1e-12)) * mask (nlevel_mg_cuda.py:262-267)
__global__ __launch_bounds__(kBlock) void inv_diag_kernel(const float* __restrict__ diagL,
                                                          const float* __restrict__ mask, float dt, float omega,
                                                          float* __restrict__ invD, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n) return;
  float d = 1.0f + dt * diagL[idx];
  d = d < 1e-12f ? 1e-12f : d;
  const float r = 1.0f / d;
  invD[idx] = (r * omega) * mask[idx];
}

// Face form of one level's coefficients (fc_common.hpp face_hi/face_lo).
__global__ __launch_bounds__(kBlock) void build_faces_kernel(Coef c, Level g, float* __restrict__ fx,
                                                             float* __restrict__ fy, float* __restrict__ fz) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= g.nx * g.ny * g.nz) return;
  int i, j, k;
  unflat(g, idx, i, j, k);
  const int sx = g.ny * g.nz;
  fx[idx] = i < g.nx - 1 ? c.aE[idx] : (lo_ghost(g, 0) ? 0.f : c.aW[idx - (g.nx - 1) * sx]);
  fy[idx] = j < g.ny - 1 ? c.aN[idx] : (lo_ghost(g, 1) ? 0.f : c.aS[idx - (g.ny - 1) * g.nz]);
  fz[idx] = k < g.nz - 1 ? c.aU[idx] : (lo_ghost(g, 2) ? 0.f : c.aD[idx - (g.nz - 1)]);
}

// Count the cells whose six coefficients or diagL, read back through the face
// form, differ in any bit from the per-cell arrays.
__global__ __launch_bounds__(kBlock) void check_faces_kernel(Coef c, Level g, unsigned* __restrict__ mismatches) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= g.nx * g.ny * g.nz) return;
  int i, j, k;
  unflat(g, idx, i, j, k);
  const C7 f = coef_at<true>(c, g, idx, i, j, k);
  const C7 a = coef_at<false>(c, g, idx, i, j, k);
  const bool same = __float_as_uint(f.e) == __float_as_uint(a.e) && __float_as_uint(f.w) == __float_as_uint(a.w) &&
                    __float_as_uint(f.n) == __float_as_uint(a.n) && __float_as_uint(f.s) == __float_as_uint(a.s) &&
                    __float_as_uint(f.u) == __float_as_uint(a.u) && __float_as_uint(f.d) == __float_as_uint(a.d) &&
                    __float_as_uint(f.diag) == __float_as_uint(a.diag);
  if (!same) atomicAdd(mismatches, 1u);
}

// 2x2x2 means of up to 6 fields (hierarchy build); mask_mode 1 thresholds
// the mean at 0.999 (nlevel_mg_cuda.py:255).
__global__ __launch_bounds__(kBlock) void restrict_field_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                                int nx, int ny, int nz, int cx, int cy, int cz,
                                                                int threshold) {
  const int64_t cidx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (cidx >= (int64_t)cx * cy * cz) return;
  const int I = (int)(cidx / ((int64_t)cy * cz)), J = (int)((cidx / cz) % cy), K = (int)(cidx % cz);
  float s = 0.f;
  for (int d = 0; d < 8; ++d) {
    const int i = 2 * I + (d >> 2), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d & 1);
    s += src[((int64_t)i * ny + j) * nz + k];
  }
  const float m = s / 8.f;
  // threshold 1: the reference's rule (unknown where every child is); 2: unknown where
  // any child is (for systems whose masked cells are decoupled by zero conductance)
  dst[cidx] = threshold == 1 ? (m >= 0.999f ? 1.f : 0.f) : (threshold == 2 ? (m > 0.f ? 1.f : 0.f) : m);
}
#pragma clang fp contract(on)

// ---- FP64 outer loop ---------------------------------------------------------
// block sum of up to 3 values into partial[blockIdx][0..2]
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64][3];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const double s = wave_sum_dpp(v[a]);
    if (lane == 0) sh[wid][a] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += sh[w][a];
      partial[(int64_t)blockIdx.x * 3 + a] = s;
    }
  }
}

// Device scalar slots used by the FCG kernels (FP64):
// S_PAP/S_PR (A p, one fold), S_RR/S_BB (true residual, one fold), S_ZAP, alpha, beta.
enum { S_PAP = 0, S_PR = 1, S_RR = 2, S_BB = 3, S_ZAP = 4, S_ALPHA = 5, S_BETA = 6, S_STOP = 7, S_N = 8 };

// fold the partials of `nparts` blocks into scal[slot0..slot0+nv-1] (one block)
__global__ __launch_bounds__(kBlock) void fold_kernel(const double* __restrict__ partial, int nparts, int nv,
                                                      double* __restrict__ scal, int slot0, int update) {
  // Each thread adds its partials p = tid, tid + kBlock, ... in that order; the
  // loads of 8 of them are issued together (one after another they would cost a
  // memory latency each).
  double v[3] = {0.0, 0.0, 0.0};
  for (int p0 = threadIdx.x; p0 < nparts; p0 += 8 * kBlock) {
    double t[8][3];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int p = p0 + u * kBlock;
#pragma unroll
      for (int a = 0; a < 3; ++a) t[u][a] = (p < nparts && a < nv) ? partial[(int64_t)p * 3 + a] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (p0 + u * kBlock < nparts) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
          if (a < nv) v[a] += t[u][a];
      }
    }
  }
  __shared__ double sh[kBlock / 64][3];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int a = 0; a < 3; ++a) {
    const double s = wave_sum_dpp(v[a]);
    if (lane == 0) sh[wid][a] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < nv; ++a) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += sh[w][a];
      scal[slot0 + a] = s;
    }
    if (update == 1) {  // alpha = (p.r) / (p.Ap)
      scal[S_ALPHA] = scal[S_PR] / scal[S_PAP];
    } else if (update == 2) {  // beta = (z.Ap_old) / (p_old.Ap_old); S_PAP still holds p_old.Ap_old
      scal[S_BETA] = scal[S_PAP] != 0.0 ? scal[S_ZAP] / scal[S_PAP] : 0.0;
    }
  }
}

// Ap = (p + dt L p) * mask in FP64 (FP32 coefficients), one cell per thread; block
// partials of p.Ap, p.r (folded in two stages: fold_runs_kernel, then fold_kernel)
template <bool FACE>
__global__ __launch_bounds__(kBlock) void apply_A64_kernel(const double* __restrict__ p, double* __restrict__ Ap,
                                                           const double* __restrict__ r, const float* __restrict__ mask,
                                                           Coef c, Level g, double dt, double* __restrict__ partial) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  double v[2] = {0.0, 0.0};
  int i = 0, j = 0, k = 0;
  if (idx < g.nx * g.ny * g.nz) unflat(g, idx, i, j, k);
  if (idx < g.nx * g.ny * g.nz && cell_live(c, idx)) {
    const Nbr o = neighbours(g, i, j, k);
    const C7 q = coef_at<FACE>(c, g, idx, i, j, k);
    const double pc = p[idx];
    const double L = (double)q.diag * pc - (double)q.e * p[idx + o.e] - (double)q.w * p[idx + o.w] -
                     (double)q.n * p[idx + o.n] - (double)q.s * p[idx + o.s] - (double)q.u * p[idx + o.u] -
                     (double)q.d * p[idx + o.d];
    const double a = (pc + dt * L) * (double)mask[idx];
    Ap[idx] = a;
    v[0] = pc * a;
    v[1] = r ? pc * r[idx] : 0.0;
  }
  block_partials<2>(v, partial);
}

// First stage of a two-stage fold: block b adds the partials of blocks
// [b * run, (b + 1) * run) (stride 3, in block order) into out[b * 3 + a].
constexpr int kFoldRuns = 256;
__global__ __launch_bounds__(kBlock) void fold_runs_kernel(const double* __restrict__ partial, int nparts, int run,
                                                           int nv, double* __restrict__ out) {
  const int b0 = blockIdx.x * run, b1 = min(b0 + run, nparts);
  double v[3] = {0.0, 0.0, 0.0};
  for (int q = b0 + (int)threadIdx.x; q < b1; q += kBlock)
    for (int a = 0; a < nv; ++a) v[a] += partial[(int64_t)q * 3 + a];
  __shared__ double sh[kBlock / 64][3];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int a = 0; a < 3; ++a) {
    const double s = wave_sum_dpp(v[a]);
    if (lane == 0) sh[wid][a] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int a = 0; a < nv; ++a) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += sh[w][a];
      out[(int64_t)blockIdx.x * 3 + a] = s;
    }
}

// x += alpha p; r -= alpha Ap; r32 = (float) r; partial of r.r
__global__ __launch_bounds__(kBlock) void update_xr_kernel(double* __restrict__ x, double* __restrict__ r,
                                                           float* __restrict__ r32, const double* __restrict__ p,
                                                           const double* __restrict__ Ap,
                                                           const double* __restrict__ scal, int64_t n,
                                                           const uint8_t* __restrict__ rows, unsigned row_len,
                                                           double* __restrict__ partial) {
  double v[1] = {0.0};
  const double alpha = scal[S_ALPHA];
  // a stopped solve (scal[S_STOP] set) keeps x and r and re-forms the same r.r
  const bool stop = scal[S_STOP] != 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    if (rows && !rows[(unsigned)idx / row_len]) continue;  // n < 2^31 (checked by the entry point)
    if (stop) {
      const double rn = r[idx];
      v[0] += rn * rn;
      continue;
    }
    x[idx] += alpha * p[idx];
    const double rn = r[idx] - alpha * Ap[idx];
    r[idx] = rn;
    r32[idx] = (float)rn;
    v[0] += rn * rn;
  }
  block_partials<1>(v, partial);
}

// partial of z.Ap_old (z is the FP32 V-cycle output)
__global__ __launch_bounds__(kBlock) void dot_z_kernel(const float* __restrict__ z, const double* __restrict__ Ap_old,
                                                       int64_t n, const uint8_t* __restrict__ rows,
                                                       unsigned row_len, double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock)
    if (!rows || rows[(unsigned)idx / row_len]) v[0] += (double)z[idx] * Ap_old[idx];
  block_partials<1>(v, partial);
}

// p = z - beta p_old (in place on p); p is kept once the solve has stopped
__global__ __launch_bounds__(kBlock) void new_dir_kernel(double* __restrict__ p, const float* __restrict__ z,
                                                         const double* __restrict__ scal, int64_t n, int first,
                                                         const uint8_t* __restrict__ rows, unsigned row_len) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n || scal[S_STOP] != 0.0) return;
  if (rows && !rows[(unsigned)idx / row_len]) return;
  p[idx] = first ? (double)z[idx] : (double)z[idx] - scal[S_BETA] * p[idx];
}

// Stopping test on the device: scal[stop_slot] becomes 1 once sqrt(scal[rr_slot]) / bnorm
// is below tol or not finite, and stays 1 until cleared. The host can then queue the
// next iteration before it reads this one's residual: the kernels of that iteration
// which would change x, r or the search directions see the flag and keep them.
__global__ void stop_check_kernel(double* __restrict__ scal, int rr_slot, int stop_slot, double bnorm, double tol,
                                  int clear) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (clear) {
    scal[stop_slot] = 0.0;
    return;
  }
  if (scal[stop_slot] != 0.0) return;
  const double rec = sqrt(scal[rr_slot]) / bnorm;
  scal[stop_slot] = (!(rec < INFINITY) || rec < tol) ? 1.0 : 0.0;
}

// true residual r = (b - A x) * mask in FP64, r32 copy, partials of r.r
template <bool FACE>
__global__ __launch_bounds__(kBlock) void true_residual_kernel(const double* __restrict__ x,
                                                               const double* __restrict__ b, double* __restrict__ r,
                                                               float* __restrict__ r32,
                                                               const float* __restrict__ mask, Coef c, Level g,
                                                               double dt, double* __restrict__ partial) {
  const int n = g.nx * g.ny * g.nz;
  double v[2] = {0.0, 0.0};
  // unrolled so that several cells' gathers are in flight; each thread still adds its
  // cells in index order
#pragma unroll 4
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < n; idx += gridDim.x * kBlock) {
    if (!cell_live(c, idx)) continue;
    int i, j, k;
    unflat(g, idx, i, j, k);
    const Nbr o = neighbours(g, i, j, k);
    const C7 q = coef_at<FACE>(c, g, idx, i, j, k);
    const double xc = x[idx];
    const double L = (double)q.diag * xc - (double)q.e * x[idx + o.e] - (double)q.w * x[idx + o.w] -
                     (double)q.n * x[idx + o.n] - (double)q.s * x[idx + o.s] - (double)q.u * x[idx + o.u] -
                     (double)q.d * x[idx + o.d];
    const double m = (double)mask[idx];
    const double bb = b[idx] * m;
    const double rr = (bb - (xc + dt * L)) * m;
    r[idx] = rr;
    r32[idx] = (float)rr;
    v[0] += rr * rr;
    v[1] += bb * bb;
  }
  block_partials<2>(v, partial);
}

// partial of sum(x^2) for an FP32 vector, FP64 accumulation
__global__ __launch_bounds__(kBlock) void norm2_f32_kernel(const float* __restrict__ x, int64_t n,
                                                           double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    const double a = (double)x[idx];
    v[0] += a * a;
  }
  block_partials<1>(v, partial);
}

// r = (b - Lo) * mask, partial of r.r (the residual of one V-cycle iteration)
__global__ __launch_bounds__(kBlock) void residual_f32_kernel(const float* __restrict__ b, const float* __restrict__ Lo,
                                                              const float* __restrict__ mask, float* __restrict__ r,
                                                              int64_t n, double* __restrict__ partial) {
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    const float a = mask ? (b[idx] - Lo[idx]) * mask[idx] : b[idx] - Lo[idx];
    r[idx] = a;
    v[0] += (double)a * (double)a;
  }
  block_partials<1>(v, partial);
}

// ---- Chebyshev smoother (chebyshev_smooth, mg_pcg.py:195-247) ----------------
// lambda_max = max(diagL * dt + 1): per-block maxima, then one folding block
__global__ __launch_bounds__(kBlock) void cheb_lam_kernel(const float* __restrict__ diagL, int64_t n, float dt,
                                                          float* __restrict__ partial) {
  __shared__ float sh[kBlock / 64];
  float m = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    m = fmaxf(m, diagL[i] * dt + 1.0f);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = sh[0];
    for (int w = 1; w < kBlock / 64; ++w) a = fmaxf(a, sh[w]);
    partial[blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(kBlock) void max_fold_kernel(const float* __restrict__ partial, int nparts,
                                                          float* __restrict__ out) {
  __shared__ float sh[kBlock / 64];
  float m = -INFINITY;
  for (int p = threadIdx.x; p < nparts; p += kBlock) m = fmaxf(m, partial[p]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = sh[0];
    for (int w = 1; w < kBlock / 64; ++w) a = fmaxf(a, sh[w]);
    out[0] = a;
  }
}

// first step: r = (b - Lo) mask, p = invD r, x_new = x + p
__global__ __launch_bounds__(kBlock) void cheb_first_kernel(const float* __restrict__ x, const float* __restrict__ b,
                                                            const float* __restrict__ Lo,
                                                            const float* __restrict__ invD,
                                                            const float* __restrict__ mask, float* __restrict__ p,
                                                            float* __restrict__ x_new, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float r = (b[i] - Lo[i]) * mask[i];
  const float pv = invD[i] * r;
  p[i] = pv;
  x_new[i] = x[i] + pv;
}

// later steps: p = rho_new (s (invD r) + rho_prev p), x_new += p; the last step
// writes x = x_new mask
__global__ __launch_bounds__(kBlock) void cheb_step_kernel(float* __restrict__ x_new, const float* __restrict__ b,
                                                           const float* __restrict__ Lo,
                                                           const float* __restrict__ invD,
                                                           const float* __restrict__ mask, float* __restrict__ p,
                                                           float rho_new, float s, float rho_prev,
                                                           float* __restrict__ x_final, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float m = mask[i];
  const float r = (b[i] - Lo[i]) * m;
  const float pn = rho_new * (s * (invD[i] * r) + rho_prev * p[i]);
  const float xn = x_new[i] + pn;
  p[i] = pn;
  x_new[i] = xn;
  if (x_final) x_final[i] = xn * m;
}

int make_level(Level& g, int nx, int ny, int nz, int ax, int out_idx) {
  FC_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "mg: every extent must be >= 2 (got %d,%d,%d)", nx, ny, nz);
  // 32-bit cell indices, including the stride of the capped reduction grids
  FC_REQUIRE((int64_t)nx * ny * nz + (int64_t)kMaxPartials * kBlock < ((int64_t)1 << 31),
             "mg: %d x %d x %d cells exceed the 32-bit index range", nx, ny, nz);
  FC_REQUIRE(ax >= 0 && ax <= 2, "mg: flow axis must be 0, 1 or 2");
  g.nx = nx; g.ny = ny; g.nz = nz; g.ax = ax; g.out_idx = out_idx;
  return 0;
}

}  // namespace

extern "C" {

struct fc_mg_coef {
  const float *aE, *aW, *aN, *aS, *aU, *aD, *diagL, *invD;
  const float *fx, *fy, *fz;
  const uint8_t* rows;
};

static Coef to_coef(const fc_mg_coef* c) {
  Coef o;
  o.aE = c->aE; o.aW = c->aW; o.aN = c->aN; o.aS = c->aS; o.aU = c->aU; o.aD = c->aD; o.diagL = c->diagL;
  o.invD = c->invD;
  o.fx = c->fx; o.fy = c->fy; o.fz = c->fz;
  o.rows = c->rows;
  return o;
}

// 1: face form, 0: per-cell arrays, -1: partly set (an error)
static int face_form(const fc_mg_coef* c) {
  const int set = (c->fx != nullptr) + (c->fy != nullptr) + (c->fz != nullptr);
  return set == 3 ? 1 : (set == 0 ? 0 : -1);
}
#define FC_COEF_FORM(c, what)                                                                  \
  const int form = face_form(c);                                                               \
  FC_REQUIRE(form >= 0, what ": face coefficients must be all set or all null");               \
  FC_REQUIRE(form == 1 || ((c)->aE && (c)->aW && (c)->aN && (c)->aS && (c)->aU && (c)->aD &&   \
                           (c)->diagL), what ": null coefficient array")

int fc_mg_mop_fused(const float* T, const float* aE, const float* aW, const float* aN, const float* aS,
                    const float* aU, const float* aD, const float* diagL, float* Lo, double dt, int nx, int ny, int nz,
                    int ax, int out_idx, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(T && aE && aW && aN && aS && aU && aD && diagL && Lo, "mop_fused: null pointer argument");
  Coef c{aE, aW, aN, aS, aU, aD, diagL, nullptr};
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(mop_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, T, c, g, (float)dt, Lo);
  return check_launch("mop_fused");
}

int fc_mg_jacobi_smooth(float* x, const float* b, const float* Lo, const float* inv_diagM, int64_t n,
                        hipStream_t stream) {
  FC_REQUIRE(n >= 0, "jacobi_smooth: n must be >= 0");
  if (n == 0) return 0;
  FC_REQUIRE(x && b && Lo && inv_diagM, "jacobi_smooth: null pointer argument");
  hipLaunchKernelGGL(jacobi_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x, b, Lo, inv_diagM, n);
  return check_launch("jacobi_smooth");
}

int fc_mg_build_coeffs(const float* k, const float* rho_cp, const float* u, const float* v, const float* w, int nx,
                       int ny, int nz, int ax, int out_idx, double dx, double s, float* aE, float* aW, float* aN,
                       float* aS, float* aU, float* aD, float* diagL, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(k && rho_cp && u && v && w && aE && aW && aN && aS && aU && aD && diagL,
             "build_coeffs: null pointer argument");
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(build_coeffs_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, k, rho_cp, u, v, w, g,
                     (float)(s / (dx * dx)), (float)(s / dx), aE, aW, aN, aS, aU, aD, diagL);
  return check_launch("build_coeffs");
}

int fc_mg_build_faces(const fc_mg_coef* coef, int nx, int ny, int nz, int ax, int out_idx, float* fx, float* fy,
                      float* fz, unsigned* mismatches, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(coef && coef->aE && coef->aW && coef->aN && coef->aS && coef->aU && coef->aD && coef->diagL,
             "build_faces: the per-cell coefficient arrays are required");
  FC_REQUIRE(fx && fy && fz && mismatches, "build_faces: null pointer argument");
  Coef c = to_coef(coef);
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(build_faces_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, c, g, fx, fy, fz);
  if (int rc = check_launch("build_faces")) return rc;
  c.fx = fx; c.fy = fy; c.fz = fz;
  hipLaunchKernelGGL(check_faces_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, c, g, mismatches);
  return check_launch("build_faces/check");
}

int fc_mg_inv_diag(const float* diagL, const float* mask, double dt, double omega, float* inv_diagM, int64_t n,
                   hipStream_t stream) {
  if (n == 0) return 0;
  FC_REQUIRE(diagL && mask && inv_diagM, "inv_diag: null pointer argument");
  hipLaunchKernelGGL(inv_diag_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, diagL, mask, (float)dt,
                     (float)omega, inv_diagM, n);
  return check_launch("inv_diag");
}

int fc_mg_restrict_field(const float* src, float* dst, int nx, int ny, int nz, int threshold, hipStream_t stream) {
  FC_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "restrict_field: extents must be >= 2");
  FC_REQUIRE(src && dst, "restrict_field: null pointer argument");
  FC_REQUIRE(threshold >= 0 && threshold <= 2, "restrict_field: threshold must be 0, 1 or 2");
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  hipLaunchKernelGGL(restrict_field_kernel, dim3(grid_for((int64_t)cx * cy * cz, kBlock)), dim3(kBlock), 0, stream,
                     src, dst, nx, ny, nz, cx, cy, cz, threshold);
  return check_launch("restrict_field");
}

int fc_mg_sweep(const float* x, float* x_out, const float* b, const fc_mg_coef* coef, double dt, int nx, int ny,
                int nz, int ax, int out_idx, int zero_guess, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(coef && coef->invD && x && x_out && b && x != x_out, "sweep: bad arguments (x and x_out must differ)");
  FC_COEF_FORM(coef, "sweep");
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(form ? sweep_kernel<true> : sweep_kernel<false>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0,
                     stream, x, x_out, b, to_coef(coef), g, (float)dt, zero_guess);
  return check_launch("sweep");
}

int fc_mg_sweep2_from_zero(float* x_out, const float* b, const fc_mg_coef* coef, double dt, int nx, int ny, int nz,
                           int ax, int out_idx, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(coef && coef->invD && x_out && b && x_out != b, "sweep2_from_zero: bad arguments");
  FC_COEF_FORM(coef, "sweep2_from_zero");
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(form ? sweep2_from_zero_kernel<true> : sweep2_from_zero_kernel<false>,
                     dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x_out, b, to_coef(coef), g, (float)dt);
  return check_launch("sweep2_from_zero");
}

int fc_mg_residual_restrict(const float* x, const float* b, const float* mask, const fc_mg_coef* coef, double dt,
                            int nx, int ny, int nz, int ax, int out_idx, float* b_coarse, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(coef && x && b && mask && b_coarse, "residual_restrict: null pointer argument");
  FC_COEF_FORM(coef, "residual_restrict");
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  FC_REQUIRE(cx >= 1 && cy >= 1 && cz >= 1, "residual_restrict: every extent must be >= 2");
  const int kchunks = (cz + 31) / 32;
  const int64_t blocks = (int64_t)cx * cy * kchunks;
  hipLaunchKernelGGL(form ? residual_restrict_kernel<true> : residual_restrict_kernel<false>, dim3((unsigned)blocks),
                     dim3(kBlock), 0, stream, x, b, mask, to_coef(coef), g, (float)dt, b_coarse, cx, cy, cz, kchunks);
  return check_launch("residual_restrict");
}

int fc_mg_prolong_correct(const float* x_coarse, float* x, const float* mask, int nx, int ny, int nz, int cx, int cy,
                          int cz, float alpha, const uint8_t* rows, hipStream_t stream) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cx > 0 && cy > 0 && cz > 0, "prolong_correct: extents must be positive");
  FC_REQUIRE(x_coarse && x && mask, "prolong_correct: null pointer argument");
  const int64_t n = (int64_t)nx * ny * nz;
  FC_REQUIRE(n < ((int64_t)1 << 31), "prolong_correct: %lld cells exceed the 32-bit index range", (long long)n);
  // the interpolation scale of F.interpolate(align_corners=False): n_in / n_out in FP32
  const float sx = (float)cx / (float)nx, sy = (float)cy / (float)ny, sz = (float)cz / (float)nz;
  if (nz % 2 == 0 && (reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(mask)) % 8 == 0)
    hipLaunchKernelGGL(prolong_correct2_kernel, dim3(grid_for(n / 2, kBlock)), dim3(kBlock), 0, stream, x_coarse, x,
                       mask, nx, ny, nz, cx, cy, cz, sx, sy, sz, alpha, rows);
  else
    hipLaunchKernelGGL(prolong_correct_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x_coarse, x, mask,
                       nx, ny, nz, cx, cy, cz, sx, sy, sz, alpha, rows);
  return check_launch("prolong_correct");
}

int fc_mg_coarse_solve(float* x, float* tmp, const float* b, const fc_mg_coef* coef, double dt, int nx, int ny,
                       int nz, int ax, int out_idx, int sweeps, int zero_guess, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(coef && coef->invD && x && tmp && b && x != tmp, "coarse_solve: bad arguments");
  FC_REQUIRE(sweeps >= 0, "coarse_solve: sweeps must be >= 0");
  FC_COEF_FORM(coef, "coarse_solve");
  if ((int64_t)nx * ny * nz <= kCoarseLdsCells)
    hipLaunchKernelGGL(form ? coarse_solve_lds_kernel<true> : coarse_solve_lds_kernel<false>, dim3(1),
                       dim3(kCoarseThreads), 0, stream, x, b, to_coef(coef), g, (float)dt, sweeps, zero_guess);
  else
    hipLaunchKernelGGL(form ? coarse_solve_kernel<true> : coarse_solve_kernel<false>, dim3(1), dim3(1024), 0, stream,
                       x, tmp, b, to_coef(coef), g, (float)dt, sweeps, zero_guess);
  return check_launch("coarse_solve");
}

// ---- FP64 flexible-CG kernels -------------------------------------------------
static unsigned reduce_grid(int64_t n) {
  const unsigned g = grid_for(n, kBlock);
  return g < kMaxPartials ? g : kMaxPartials;
}

// block partials (3 doubles each) for the reductions: the capped grids, or one block per
// 256 cells (apply_A64) followed by the runs of its two-stage fold
int fc_mg_partials_count(int64_t n) {
  const int64_t g = grid_for(n, kBlock) + kFoldRuns;
  return (int)(g > (int64_t)reduce_grid(n) ? g : (int64_t)reduce_grid(n));
}

int fc_mg_apply_A64(const double* p, double* Ap, const double* r, const float* mask, const fc_mg_coef* coef,
                    double dt, int nx, int ny, int nz, int ax, int out_idx, double* partial, double* scal,
                    hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(p && Ap && mask && coef && partial && scal, "apply_A64: null pointer argument");
  FC_COEF_FORM(coef, "apply_A64");
  const int64_t n = (int64_t)nx * ny * nz;
  const int nb = (int)grid_for(n, kBlock);
  hipLaunchKernelGGL(form ? apply_A64_kernel<true> : apply_A64_kernel<false>, dim3(nb), dim3(kBlock), 0, stream, p,
                     Ap, r, mask, to_coef(coef), g, dt, partial);
  if (int rc = check_launch("apply_A64")) return rc;
  const int runs = nb < kFoldRuns ? nb : kFoldRuns, run = (nb + runs - 1) / runs;
  double* second = partial + (int64_t)nb * 3;  // within fc_mg_partials_count(n) * 3 doubles
  hipLaunchKernelGGL(fold_runs_kernel, dim3(runs), dim3(kBlock), 0, stream, (const double*)partial, nb, run, 2,
                     second);
  if (int rc = check_launch("apply_A64/fold runs")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const double*)second, runs, 2, scal, (int)S_PAP,
                     r ? 1 : 0);
  return check_launch("apply_A64/fold");
}

int fc_mg_update_xr(double* x, double* r, float* r32, const double* p, const double* Ap, int64_t n,
                    const uint8_t* rows, int row_len, double* partial, double* scal, hipStream_t stream) {
  FC_REQUIRE(x && r && r32 && p && Ap && partial && scal, "update_xr: null pointer argument");
  FC_REQUIRE(!rows || (row_len > 0 && n < ((int64_t)1 << 31)),
             "update_xr: rows need a positive run length and n < 2^31");
  const unsigned gsz = reduce_grid(n);
  hipLaunchKernelGGL(update_xr_kernel, dim3(gsz), dim3(kBlock), 0, stream, x, r, r32, p, Ap, scal, n, rows,
                     (unsigned)(rows ? row_len : 1), partial);
  if (int rc = check_launch("update_xr")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, partial, (int)gsz, 1, scal, (int)S_RR, 0);
  return check_launch("update_xr/fold");
}

int fc_solver_stop_check(double* scal, int rr_slot, int stop_slot, double bnorm, double tol, int clear,
                         hipStream_t stream) {
  FC_REQUIRE(scal, "stop_check: null scalar array");
  FC_REQUIRE(rr_slot >= 0 && stop_slot >= 0 && rr_slot != stop_slot, "stop_check: bad slots");
  FC_REQUIRE(bnorm > 0.0, "stop_check: ||b|| must be positive");
  hipLaunchKernelGGL(stop_check_kernel, dim3(1), dim3(64), 0, stream, scal, rr_slot, stop_slot, bnorm, tol, clear);
  return check_launch("stop_check");
}

int fc_mg_stop_slot(void) { return S_STOP; }

int fc_mg_new_direction(double* p, const float* z, const double* Ap_old, int64_t n, int first, const uint8_t* rows,
                        int row_len, double* partial, double* scal, hipStream_t stream) {
  FC_REQUIRE(p && z && partial && scal, "new_direction: null pointer argument");
  FC_REQUIRE(!rows || (row_len > 0 && n < ((int64_t)1 << 31)),
             "new_direction: rows need a positive run length and n < 2^31");
  const unsigned rl = (unsigned)(rows ? row_len : 1);
  const unsigned gsz = reduce_grid(n);
  if (!first) {
    FC_REQUIRE(Ap_old, "new_direction: Ap_old required after the first iteration");
    hipLaunchKernelGGL(dot_z_kernel, dim3(gsz), dim3(kBlock), 0, stream, z, Ap_old, n, rows, rl, partial);
    if (int rc = check_launch("new_direction/dot")) return rc;
    hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, partial, (int)gsz, 1, scal, (int)S_ZAP, 2);
    if (int rc = check_launch("new_direction/fold")) return rc;
  }
  hipLaunchKernelGGL(new_dir_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, p, z, scal, n, first, rows,
                     rl);
  return check_launch("new_direction");
}

int fc_mg_true_residual(const double* x, const double* b, double* r, float* r32, const float* mask,
                        const fc_mg_coef* coef, double dt, int nx, int ny, int nz, int ax, int out_idx,
                        double* partial, double* scal, hipStream_t stream) {
  Level g;
  if (int rc = make_level(g, nx, ny, nz, ax, out_idx)) return rc;
  FC_REQUIRE(x && b && r && r32 && mask && coef && partial && scal, "true_residual: null pointer argument");
  FC_COEF_FORM(coef, "true_residual");
  const int64_t n = (int64_t)nx * ny * nz;
  const unsigned gsz = reduce_grid(n);
  hipLaunchKernelGGL(form ? true_residual_kernel<true> : true_residual_kernel<false>, dim3(gsz), dim3(kBlock), 0, stream, x, b, r, r32, mask, to_coef(coef), g,
                     dt, partial);
  if (int rc = check_launch("true_residual")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, partial, (int)gsz, 2, scal, (int)S_RR, 0);
  return check_launch("true_residual/fold");
}

int fc_mg_norm2_f32(const float* x, int64_t n, double* partial, double* scal, int slot, hipStream_t stream) {
  FC_REQUIRE(x && partial && scal && slot >= 0 && slot < S_N, "norm2_f32: bad arguments");
  if (n == 0) return 0;
  const unsigned gsz = reduce_grid(n);
  hipLaunchKernelGGL(norm2_f32_kernel, dim3(gsz), dim3(kBlock), 0, stream, x, n, partial);
  if (int rc = check_launch("norm2_f32")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, partial, (int)gsz, 1, scal, slot, 0);
  return check_launch("norm2_f32/fold");
}

int fc_mg_residual_f32(const float* b, const float* Lo, const float* mask, float* r, int64_t n, double* partial,
                       double* scal, int slot, hipStream_t stream) {
  FC_REQUIRE(b && Lo && r && partial && scal && slot >= 0 && slot < S_N, "residual_f32: bad arguments");
  if (n == 0) return 0;
  const unsigned gsz = reduce_grid(n);
  hipLaunchKernelGGL(residual_f32_kernel, dim3(gsz), dim3(kBlock), 0, stream, b, Lo, mask, r, n, partial);
  if (int rc = check_launch("residual_f32")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, partial, (int)gsz, 1, scal, slot, 0);
  return check_launch("residual_f32/fold");
}

int fc_mg_scalar_slots(void) { return S_N; }

int fc_mg_cheb_lambda_max(const float* diagL, int64_t n, double dt, float* partial, float* out, hipStream_t stream) {
  FC_REQUIRE(diagL && partial && out && n > 0, "cheb_lambda_max: bad arguments");
  const unsigned gsz = reduce_grid(n);
  hipLaunchKernelGGL(cheb_lam_kernel, dim3(gsz), dim3(kBlock), 0, stream, diagL, n, (float)dt, partial);
  if (int rc = check_launch("cheb_lambda_max")) return rc;
  hipLaunchKernelGGL(max_fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const float*)partial, (int)gsz, out);
  return check_launch("cheb_lambda_max/fold");
}

int fc_mg_cheb_first(const float* x, const float* b, const float* Lo, const float* invD, const float* mask, float* p,
                     float* x_new, int64_t n, hipStream_t stream) {
  FC_REQUIRE(x && b && Lo && invD && mask && p && x_new && n >= 0, "cheb_first: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(cheb_first_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x, b, Lo, invD, mask, p,
                     x_new, n);
  return check_launch("cheb_first");
}

int fc_mg_cheb_step(float* x_new, const float* b, const float* Lo, const float* invD, const float* mask, float* p,
                    double rho_new, double s, double rho_prev, float* x_final, int64_t n, hipStream_t stream) {
  FC_REQUIRE(x_new && b && Lo && invD && mask && p && n >= 0, "cheb_step: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(cheb_step_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x_new, b, Lo, invD, mask,
                     p, (float)rho_new, (float)s, (float)rho_prev, x_final, n);
  return check_launch("cheb_step");
}

int fc_mg_cheb_partials_count(int64_t n) { return (int)reduce_grid(n); }

}  // extern "C"
