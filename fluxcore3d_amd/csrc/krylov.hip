// krylov.hip — flexible GCR(m) in FP64 for the conduction system A x = b.
//
// Minimal-residual Krylov iteration for a non-symmetric operator and a
// variable preconditioner (the learned one): with z = P(r),
//   p = z,  w = A p,
//   w -= sum_j c_j W_j,  p -= sum_j c_j P_j,  c_j = (w . W_j) / (W_j . W_j)   (last m pairs)
//   alpha = (r . w) / (w . w),  x += alpha p,  r -= alpha w.
// Each step removes the component of r along the new, orthogonalised A-image, so
// ||r|| never grows. Vectors are FP64, the operator uses the FP32 coefficients
// of the finest multigrid level; scalars stay in device memory. Reductions run
// a fixed grid (one partial per block, folded by one workgroup), so the
// summation order is fixed.
#include "fc_common.hpp"

using namespace fc;

namespace {

constexpr int kBlock = 256;
// grid cap of the GCR reductions (16384 and 32768 made the 256^3 GCR(6) solve 1.5 % and
// 6 % slower: their single-workgroup folds grow with the grid)
constexpr unsigned kMaxBlocks = 8192;
constexpr int kMaxDirs = 8;

// scalar slots (per call family)
enum { K_RW = 0, K_WW = 1, K_RR = 2, K_ALPHA = 3, K_STOP = 7, K_DOT0 = 8, K_WWHIST0 = 16, K_SLOTS = 24 };

struct Stencil {
  const float *aE, *aW, *aN, *aS, *aU, *aD, *diagL;
  const float* mask;
  int nx, ny, nz, ax, out_idx;
  double dt;
  const float *fx, *fy, *fz;  // face form of the coefficients (fc_common.hpp), or null
};

struct Dirs {
  const double* W[kMaxDirs];
  double* P[kMaxDirs];
  int slot[kMaxDirs];  // history slot of each direction's W.W
  int m;
};

__device__ __forceinline__ int lo_ghost(const Stencil& s, int axis) {
  return (s.ax != axis || s.out_idx == 0) ? 1 : 0;
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partial, int stride) {
  __shared__ double sh[kBlock / 64][NV];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const double s = wave_sum_dpp(v[a]);
    if (lane == 0) sh[wid][a] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w][threadIdx.x];
    partial[(int64_t)blockIdx.x * stride + threadIdx.x] = s;
  }
}

// scal[dst + a] = sum over blocks of partial[b * stride + a], a < nv. Each thread
// adds its blocks b = tid, tid + kBlock, ... in that order, the loads of 4 blocks
// (all nv values) issued together.
__global__ __launch_bounds__(kBlock) void fold_kernel(const double* __restrict__ partial, int nblocks, int stride,
                                                      int nv, double* __restrict__ scal, int dst) {
  double v[kMaxDirs];
#pragma unroll
  for (int a = 0; a < kMaxDirs; ++a) v[a] = 0.0;
  for (int b0 = threadIdx.x; b0 < nblocks; b0 += 4 * kBlock) {
    double t[4][kMaxDirs];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kBlock;
#pragma unroll
      for (int a = 0; a < kMaxDirs; ++a) t[u][a] = (b < nblocks && a < nv) ? partial[(int64_t)b * stride + a] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (b0 + u * kBlock < nblocks) {
#pragma unroll
        for (int a = 0; a < kMaxDirs; ++a)
          if (a < nv) v[a] += t[u][a];
      }
    }
  }
  __shared__ double sh[kBlock / 64][kMaxDirs];
#pragma unroll
  for (int a = 0; a < kMaxDirs; ++a) {
    const double w = wave_sum_dpp(v[a]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][a] = w;
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w][threadIdx.x];
    scal[dst + threadIdx.x] = s;
  }
}

// p = (double) z;  w = (p + dt L p) * mask  (one cell per thread, 32-bit index math).
// FACE: the six coefficients come from the face arrays and diagL is their sum in
// the coefficient build's order (the same bits as the per-cell arrays).
template <bool FACE>
__global__ __launch_bounds__(kBlock) void apply_kernel(const float* __restrict__ z, Stencil s, double* __restrict__ p,
                                                       double* __restrict__ w) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= s.nx * s.ny * s.nz) return;
  const unsigned nynz = (unsigned)s.ny * (unsigned)s.nz, u = (unsigned)idx;
  const int i = (int)(u / nynz);
  const unsigned rem = u - (unsigned)i * nynz;
  const int j = (int)(rem / (unsigned)s.nz), k = (int)(rem - (unsigned)j * (unsigned)s.nz);
  // neighbour offsets with the reference's ghosts (mg_kernels.cu:37-44)
  const int oE = (i < s.nx - 1 ? 1 : -1) * (int)nynz, oW = (i > 0 ? -1 : lo_ghost(s, 0)) * (int)nynz;
  const int oN = (j < s.ny - 1 ? 1 : -1) * s.nz, oS = (j > 0 ? -1 : lo_ghost(s, 1)) * s.nz;
  const int oU = k < s.nz - 1 ? 1 : -1, oD = k > 0 ? -1 : lo_ghost(s, 2);
  float cE, cW, cN, cS, cU, cD, cL;
  if constexpr (FACE) {
    cE = s.fx[face_hi(idx, i, s.nx, (int)nynz)];
    cW = s.fx[face_lo(idx, i, s.nx, (int)nynz, lo_ghost(s, 0))];
    cN = s.fy[face_hi(idx, j, s.ny, s.nz)];
    cS = s.fy[face_lo(idx, j, s.ny, s.nz, lo_ghost(s, 1))];
    cU = s.fz[face_hi(idx, k, s.nz, 1)];
    cD = s.fz[face_lo(idx, k, s.nz, 1, lo_ghost(s, 2))];
    cL = ((((cE + cW) + cN) + cS) + cU) + cD;
  } else {
    cE = s.aE[idx]; cW = s.aW[idx]; cN = s.aN[idx]; cS = s.aS[idx]; cU = s.aU[idx]; cD = s.aD[idx];
    cL = s.diagL[idx];
  }
  const double pc = (double)z[idx];
  const double L = (double)cL * pc - (double)cE * (double)z[idx + oE] - (double)cW * (double)z[idx + oW] -
                   (double)cN * (double)z[idx + oN] - (double)cS * (double)z[idx + oS] -
                   (double)cU * (double)z[idx + oU] - (double)cD * (double)z[idx + oD];
  p[idx] = pc;
  w[idx] = (pc + s.dt * L) * (double)s.mask[idx];
}

// partials of w . W_j (j < m)
__global__ __launch_bounds__(kBlock) void dots_kernel(const double* __restrict__ w, Dirs d, int64_t n,
                                                      double* __restrict__ partial) {
  double v[kMaxDirs];
#pragma unroll
  for (int j = 0; j < kMaxDirs; ++j) v[j] = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    const double wi = w[idx];
#pragma unroll
    for (int j = 0; j < kMaxDirs; ++j)
      if (j < d.m) v[j] += wi * d.W[j][idx];
  }
  block_partials<kMaxDirs>(v, partial, kMaxDirs);
}

// w -= sum c_j W_j; p -= sum c_j P_j with c_j = scal[K_DOT0 + j] / scal[slot_j]; partials of r.w, w.w
__global__ __launch_bounds__(kBlock) void orth_kernel(double* __restrict__ w, double* __restrict__ p,
                                                      const double* __restrict__ r, Dirs d,
                                                      const double* __restrict__ scal, int64_t n,
                                                      double* __restrict__ partial) {
  double c[kMaxDirs];
#pragma unroll
  for (int j = 0; j < kMaxDirs; ++j) {
    const double den = j < d.m ? scal[d.slot[j]] : 0.0;
    c[j] = (j < d.m && den != 0.0) ? scal[K_DOT0 + j] / den : 0.0;
  }
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    double wi = w[idx], pi = p[idx];
#pragma unroll
    for (int j = 0; j < kMaxDirs; ++j) {
      if (j < d.m) {
        wi -= c[j] * d.W[j][idx];
        pi -= c[j] * d.P[j][idx];
      }
    }
    w[idx] = wi;
    p[idx] = pi;
    v[0] += r[idx] * wi;
    v[1] += wi * wi;
  }
  block_partials<2>(v, partial, 2);
}

// alpha = r.w / w.w; x += alpha p; r -= alpha w; r32 = (float) r; partials of r.r
__global__ __launch_bounds__(kBlock) void update_kernel(double* __restrict__ x, double* __restrict__ r,
                                                        float* __restrict__ r32, const double* __restrict__ p,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ scal, int64_t n,
                                                        double* __restrict__ partial) {
  const double ww = scal[K_WW];
  const double alpha = ww > 0.0 ? scal[K_RW] / ww : 0.0;
  // a stopped solve (scal[K_STOP] set by fc_solver_stop_check) keeps x and r
  const bool stop = scal[K_STOP] != 0.0;
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    if (stop) {
      const double rn = r[idx];
      v[0] += rn * rn;
      continue;
    }
    x[idx] += alpha * p[idx];
    const double rn = r[idx] - alpha * w[idx];
    r[idx] = rn;
    r32[idx] = (float)rn;
    v[0] += rn * rn;
  }
  block_partials<1>(v, partial, 1);
}

// ---- restarted flexible GMRES(m) --------------------------------------------------
// The Arnoldi basis V_0..V_m is kept unnormalised in FP64 with ||V_i||^2 in device
// scalars; the preconditioner reads an FP32 copy of the newest V_j. One Arnoldi step
// is two passes: w = A z (z the FP32 preconditioned vector) written as V_{j+1},
// with the partials of w . V_i, i <= j; then V_{j+1} = w - sum_i (w.V_i / V_i.V_i) V_i
// in place (classical Gram-Schmidt), its FP32 copy and ||V_{j+1}||^2.
constexpr int kMaxBasis = 24;
// grid cap of the Arnoldi kernels (at 256^3: 2 cells per thread; 8192 / 16384 / 32768 /
// 65536 blocks gave 77.5 / 76.3 / 75.7 / 76.0 ms for the FGMRES(16) bench solve)
constexpr unsigned kArnoldiBlocks = 32768;
enum { F_DOT0 = 0, F_NV0 = kMaxBasis + 1, F_SLOTS = 2 * (kMaxBasis + 1) };

struct Basis {
  const double* V[kMaxBasis + 1];
};

// column-major partials (partial[a * nblocks + b]); block a folds column a, in
// block order, into scal[dst + a]
__global__ __launch_bounds__(kBlock) void fold_columns_kernel(const double* __restrict__ partial, int nblocks,
                                                              double* __restrict__ scal, int dst) {
  const double* col = partial + (int64_t)blockIdx.x * nblocks;
  double v = 0.0;
  for (int b0 = threadIdx.x; b0 < nblocks; b0 += 8 * kBlock) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = b0 + u * kBlock < nblocks ? col[b0 + u * kBlock] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b0 + u * kBlock < nblocks) v += t[u];
  }
  __shared__ double sh[kBlock / 64];
  v = wave_sum_dpp(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
    scal[dst + blockIdx.x] = s;
  }
}

// block sums of acc[0..NV-1] into column-major partials
template <int NV>
__device__ __forceinline__ void block_columns(const double* acc, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64][NV];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const double s = wave_sum_dpp(acc[a]);
    if (lane == 0) sh[wid][a] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w][threadIdx.x];
    partial[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// The basis count NV is a template parameter: the NV loads of a cell are then
// issued together (a runtime bound would put each behind its own branch and wait
// for it in turn).
// w = (z + dt L z) * mask in FP64 (the same value as apply_kernel's w), written to
// wout; partials of w . V_i for i < NV
template <bool FACE, int NV>
__global__ __launch_bounds__(kBlock) void arnoldi_apply_kernel(const float* __restrict__ z, Stencil s,
                                                               double* __restrict__ wout, Basis bs,
                                                               double* __restrict__ partial) {
  double acc[NV];
#pragma unroll
  for (int a = 0; a < NV; ++a) acc[a] = 0.0;
  const int n = s.nx * s.ny * s.nz;
  const unsigned nynz = (unsigned)s.ny * (unsigned)s.nz;
  for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < n; idx += gridDim.x * kBlock) {
    const unsigned u = (unsigned)idx;
    const int i = (int)(u / nynz);
    const unsigned rem = u - (unsigned)i * nynz;
    const int j = (int)(rem / (unsigned)s.nz), k = (int)(rem - (unsigned)j * (unsigned)s.nz);
    const int oE = (i < s.nx - 1 ? 1 : -1) * (int)nynz, oW = (i > 0 ? -1 : lo_ghost(s, 0)) * (int)nynz;
    const int oN = (j < s.ny - 1 ? 1 : -1) * s.nz, oS = (j > 0 ? -1 : lo_ghost(s, 1)) * s.nz;
    const int oU = k < s.nz - 1 ? 1 : -1, oD = k > 0 ? -1 : lo_ghost(s, 2);
    float cE, cW, cN, cS, cU, cD, cL;
    if constexpr (FACE) {
      cE = s.fx[face_hi(idx, i, s.nx, (int)nynz)];
      cW = s.fx[face_lo(idx, i, s.nx, (int)nynz, lo_ghost(s, 0))];
      cN = s.fy[face_hi(idx, j, s.ny, s.nz)];
      cS = s.fy[face_lo(idx, j, s.ny, s.nz, lo_ghost(s, 1))];
      cU = s.fz[face_hi(idx, k, s.nz, 1)];
      cD = s.fz[face_lo(idx, k, s.nz, 1, lo_ghost(s, 2))];
      cL = ((((cE + cW) + cN) + cS) + cU) + cD;
    } else {
      cE = s.aE[idx]; cW = s.aW[idx]; cN = s.aN[idx]; cS = s.aS[idx]; cU = s.aU[idx]; cD = s.aD[idx];
      cL = s.diagL[idx];
    }
    double vb[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) vb[a] = bs.V[a][idx];
    const double pc = (double)z[idx];
    const double L = (double)cL * pc - (double)cE * (double)z[idx + oE] - (double)cW * (double)z[idx + oW] -
                     (double)cN * (double)z[idx + oN] - (double)cS * (double)z[idx + oS] -
                     (double)cU * (double)z[idx + oU] - (double)cD * (double)z[idx + oD];
    const double wv = (pc + s.dt * L) * (double)s.mask[idx];
    wout[idx] = wv;
#pragma unroll
    for (int a = 0; a < NV; ++a) acc[a] += wv * vb[a];
  }
  block_columns<NV>(acc, partial);
}

// v = v - sum_{i<NV} (scal[F_DOT0+i] / scal[F_NV0+i]) V_i in place, v32 = (float) v;
// partials of v . v
template <int NV>
__global__ __launch_bounds__(kBlock) void arnoldi_orth_kernel(double* __restrict__ v, float* __restrict__ v32,
                                                              Basis bs, const double* __restrict__ scal, int64_t n,
                                                              double* __restrict__ partial) {
  double c[NV];
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const double den = scal[F_NV0 + a];
    c[a] = den > 0.0 ? scal[F_DOT0 + a] / den : 0.0;
  }
  double acc[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    double vb[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) vb[a] = bs.V[a][idx];
    double vi = v[idx];
#pragma unroll
    for (int a = 0; a < NV; ++a) vi -= c[a] * vb[a];
    v[idx] = vi;
    v32[idx] = (float)vi;
    acc[0] += vi * vi;
  }
  block_columns<1>(acc, partial);
}

struct ZSet {
  const float* Z[kMaxBasis];
  double a[kMaxBasis];
};

// x += sum_{i<K} a_i Z_i (the sum formed first, in i order)
template <int K>
__global__ __launch_bounds__(kBlock) void arnoldi_update_x_kernel(double* __restrict__ x, ZSet zs, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n) return;
  float zb[K];
#pragma unroll
  for (int a = 0; a < K; ++a) zb[a] = zs.Z[a][idx];
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < K; ++a) s += zs.a[a] * (double)zb[a];
  x[idx] += s;
}

// Givens state of one GMRES cycle (FP64, device): R (upper triangle, column-major with
// leading dimension kMaxBasis + 1), the rotations, g and the residual estimate.
enum { G_R = 0, G_CS = (kMaxBasis + 1) * kMaxBasis, G_SN = G_CS + kMaxBasis, G_G = G_SN + kMaxBasis,
       G_EST = G_G + kMaxBasis + 1, G_SLOTS = G_EST + 1 };

// Column j of the normalised Hessenberg matrix from the Arnoldi scalars
// (h_ij = w.V_i / (|V_j| |V_i|), h_{j+1,j} = |V_{j+1}| / |V_j|), the previous rotations
// applied, a new rotation zeroing h_{j+1,j}, g rotated; est = |g_{j+1}| / bnorm.
// j = 0 starts the cycle: g = (|V_0|, 0, ...).
__global__ void arnoldi_givens_kernel(const double* __restrict__ scal, double* __restrict__ st, int j,
                                      double bnorm) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* R = st + G_R;
  double* cs = st + G_CS;
  double* sn = st + G_SN;
  double* g = st + G_G;
  if (j == 0) {
    for (int i = 0; i <= kMaxBasis; ++i) g[i] = 0.0;
    g[0] = sqrt(scal[F_NV0]);
  }
  double* Rj = R + j * (kMaxBasis + 1);
  const double nvj = scal[F_NV0 + j];
  if (!(nvj > 0.0)) {
    // V_j = 0 (exact breakdown): step j-1 already gave the solution in this Krylov
    // space (its estimate g[j] is 0). Column j is left empty with an identity
    // rotation; the host drops it from the update.
    for (int i = 0; i <= j; ++i) Rj[i] = 0.0;
    cs[j] = 1.0;
    sn[j] = 0.0;
    g[j + 1] = 0.0;
    st[G_EST] = fabs(g[j]) / bnorm;
    return;
  }
  double col[kMaxBasis + 1];
  const double hj = sqrt(nvj);
  for (int i = 0; i <= j; ++i) {
    const double nvi = scal[F_NV0 + i];
    col[i] = nvi > 0.0 ? scal[F_DOT0 + i] / (hj * sqrt(nvi)) : 0.0;
  }
  col[j + 1] = sqrt(scal[F_NV0 + j + 1]) / hj;
  for (int i = 0; i < j; ++i) {
    const double a = col[i], b = col[i + 1];
    col[i] = cs[i] * a + sn[i] * b;
    col[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double d = hypot(col[j], col[j + 1]);
  const double c = d > 0.0 ? col[j] / d : 1.0, sv = d > 0.0 ? col[j + 1] / d : 0.0;
  cs[j] = c;
  sn[j] = sv;
  col[j] = d;
  g[j + 1] = -sv * g[j];
  g[j] = c * g[j];
  for (int i = 0; i <= j; ++i) Rj[i] = col[i];
  st[G_EST] = fabs(g[j + 1]) / bnorm;
}

// ---- MG-preconditioned BiCGSTAB (mg_pcg.py:103-175) ------------------------------
// FP64 vectors, FP32 preconditioned vectors; the scalars (alpha, beta, omega) are
// formed on the host from the dot products, in the reference's order.

// out = (z + dt L z) [* mask] in FP64 from the FP32 z (the reference's Lo of the
// preconditioned vector; it leaves the masked rows in)
template <bool FACE>
__global__ __launch_bounds__(kBlock) void krylov_apply_kernel(const float* __restrict__ z, Stencil s,
                                                              double* __restrict__ out, int use_mask) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= s.nx * s.ny * s.nz) return;
  const unsigned nynz = (unsigned)s.ny * (unsigned)s.nz, u = (unsigned)idx;
  const int i = (int)(u / nynz);
  const unsigned rem = u - (unsigned)i * nynz;
  const int j = (int)(rem / (unsigned)s.nz), k = (int)(rem - (unsigned)j * (unsigned)s.nz);
  const int oE = (i < s.nx - 1 ? 1 : -1) * (int)nynz, oW = (i > 0 ? -1 : lo_ghost(s, 0)) * (int)nynz;
  const int oN = (j < s.ny - 1 ? 1 : -1) * s.nz, oS = (j > 0 ? -1 : lo_ghost(s, 1)) * s.nz;
  const int oU = k < s.nz - 1 ? 1 : -1, oD = k > 0 ? -1 : lo_ghost(s, 2);
  float cE, cW, cN, cS, cU, cD, cL;
  if constexpr (FACE) {
    cE = s.fx[face_hi(idx, i, s.nx, (int)nynz)];
    cW = s.fx[face_lo(idx, i, s.nx, (int)nynz, lo_ghost(s, 0))];
    cN = s.fy[face_hi(idx, j, s.ny, s.nz)];
    cS = s.fy[face_lo(idx, j, s.ny, s.nz, lo_ghost(s, 1))];
    cU = s.fz[face_hi(idx, k, s.nz, 1)];
    cD = s.fz[face_lo(idx, k, s.nz, 1, lo_ghost(s, 2))];
    cL = ((((cE + cW) + cN) + cS) + cU) + cD;
  } else {
    cE = s.aE[idx]; cW = s.aW[idx]; cN = s.aN[idx]; cS = s.aS[idx]; cU = s.aU[idx]; cD = s.aD[idx];
    cL = s.diagL[idx];
  }
  const double pc = (double)z[idx];
  const double L = (double)cL * pc - (double)cE * (double)z[idx + oE] - (double)cW * (double)z[idx + oW] -
                   (double)cN * (double)z[idx + oN] - (double)cS * (double)z[idx + oS] -
                   (double)cU * (double)z[idx + oU] - (double)cD * (double)z[idx + oD];
  const double lo = pc + s.dt * L;
  out[idx] = use_mask ? lo * (double)s.mask[idx] : lo;
}

struct DotPairs {
  const double* a[3];
  const double* b[3];
};

// partials of a_i . b_i, i < NP
template <int NP>
__global__ __launch_bounds__(kBlock) void krylov_dots_kernel(DotPairs d, int64_t n, double* __restrict__ partial) {
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
#pragma unroll 2
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] += d.a[q][idx] * d.b[q][idx];
  }
  block_columns<NP>(acc, partial);
}

// p = (beta p + r) - (beta omega) v;  p32 = (float) (p * mask)
__global__ __launch_bounds__(kBlock) void bicgstab_dir_kernel(double* __restrict__ p, const double* __restrict__ r,
                                                              const double* __restrict__ v,
                                                              const float* __restrict__ mask, double beta,
                                                              double bw, float* __restrict__ p32, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n) return;
  const double pn = (beta * p[idx] + r[idx]) - bw * v[idx];
  p[idx] = pn;
  p32[idx] = (float)(pn * (double)mask[idx]);
}

// s = (r - alpha v) * mask;  s32 = (float) s
__global__ __launch_bounds__(kBlock) void bicgstab_s_kernel(double* __restrict__ sv, const double* __restrict__ r,
                                                            const double* __restrict__ v,
                                                            const float* __restrict__ mask, double alpha,
                                                            float* __restrict__ s32, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n) return;
  const double sn = (r[idx] - alpha * v[idx]) * (double)mask[idx];
  sv[idx] = sn;
  s32[idx] = (float)sn;
}

// x = ((x + alpha p_hat) + omega s_hat) * mask;  r = (s - omega t) * mask (omega_r = 0 and
// s_hat = nullptr: x = (x + alpha p_hat) * mask only, r untouched)
__global__ __launch_bounds__(kBlock) void bicgstab_x_kernel(double* __restrict__ x, double* __restrict__ r,
                                                            const double* __restrict__ sv,
                                                            const double* __restrict__ t,
                                                            const float* __restrict__ p_hat,
                                                            const float* __restrict__ s_hat,
                                                            const float* __restrict__ mask, double alpha,
                                                            double omega, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n) return;
  const double m = (double)mask[idx];
  double xv = x[idx] + alpha * (double)p_hat[idx];
  if (s_hat) {
    xv = xv + omega * (double)s_hat[idx];
    r[idx] = (sv[idx] - omega * t[idx]) * m;
  }
  x[idx] = xv * m;
}

// launches the instantiation with NV == nv (1 <= nv <= kMaxBasis)
template <int NV>
struct ArnoldiLaunch {
  static void apply(int nv, bool faces, unsigned g, hipStream_t st, const float* z, const Stencil& s, double* w,
                    const Basis& bs, double* partial) {
    if (nv != NV) return ArnoldiLaunch<NV - 1>::apply(nv, faces, g, st, z, s, w, bs, partial);
    if (faces)
      hipLaunchKernelGGL((arnoldi_apply_kernel<true, NV>), dim3(g), dim3(kBlock), 0, st, z, s, w, bs, partial);
    else
      hipLaunchKernelGGL((arnoldi_apply_kernel<false, NV>), dim3(g), dim3(kBlock), 0, st, z, s, w, bs, partial);
  }
  static void orth(int nv, unsigned g, hipStream_t st, double* v, float* v32, const Basis& bs, const double* scal,
                   int64_t n, double* partial) {
    if (nv != NV) return ArnoldiLaunch<NV - 1>::orth(nv, g, st, v, v32, bs, scal, n, partial);
    hipLaunchKernelGGL(arnoldi_orth_kernel<NV>, dim3(g), dim3(kBlock), 0, st, v, v32, bs, scal, n, partial);
  }
  static void update(int k, hipStream_t st, double* x, const ZSet& zs, int64_t n) {
    if (k != NV) return ArnoldiLaunch<NV - 1>::update(k, st, x, zs, n);
    hipLaunchKernelGGL(arnoldi_update_x_kernel<NV>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, x, zs, n);
  }
};
template <>
struct ArnoldiLaunch<0> {
  static void apply(int, bool, unsigned, hipStream_t, const float*, const Stencil&, double*, const Basis&, double*) {}
  static void orth(int, unsigned, hipStream_t, double*, float*, const Basis&, const double*, int64_t, double*) {}
  static void update(int, hipStream_t, double*, const ZSet&, int64_t) {}
};

__global__ void copy_scalar_kernel(double* __restrict__ scal, int src, int dst) { scal[dst] = scal[src]; }

unsigned grid_of(int64_t n) {
  const unsigned g = grid_for(n, kBlock);
  return g < kMaxBlocks ? g : kMaxBlocks;
}

}  // namespace

extern "C" {

struct fc_gcr_stencil {
  const float *aE, *aW, *aN, *aS, *aU, *aD, *diagL, *mask;
  double dt;
  int nx, ny, nz, ax, out_idx;
  const float *fx, *fy, *fz;
};

int fc_gcr_scalar_slots(void) { return K_SLOTS; }
int fc_gcr_stop_slot(void) { return K_STOP; }
int fc_gcr_max_directions(void) { return kMaxDirs; }
int fc_gcr_partials_count(int64_t n) { return (int)grid_of(n) * kMaxDirs; }

// One flexible GCR step after the preconditioner produced z32 (FP32):
// p, w = A p, orthogonalisation against the m stored (P_j, W_j) with their W_j.W_j
// in scal[ww_slot[j]], alpha, x/r update, ||r||^2 -> scal[K_RR], w.w -> scal[store_slot].
int fc_gcr_step(const float* z32, const fc_gcr_stencil* st, double* p, double* w, const double* const* Wd,
                double* const* Pd, const int* ww_slot, int m, int store_slot, double* x, double* r, float* r32,
                double* partial, double* scal, hipStream_t stream) {
  FC_REQUIRE(z32 && st && p && w && x && r && r32 && partial && scal, "gcr_step: null pointer argument");
  FC_REQUIRE(m >= 0 && m <= kMaxDirs, "gcr_step: 0 <= m <= %d directions", kMaxDirs);
  FC_REQUIRE(store_slot >= K_WWHIST0 && store_slot < K_SLOTS, "gcr_step: store slot out of range");
  FC_REQUIRE(st->nx >= 2 && st->ny >= 2 && st->nz >= 2, "gcr_step: every extent must be >= 2");
  Stencil s{st->aE, st->aW, st->aN, st->aS, st->aU, st->aD, st->diagL, st->mask, st->nx, st->ny, st->nz,
            st->ax, st->out_idx, st->dt, st->fx, st->fy, st->fz};
  const int faces = (s.fx != nullptr) + (s.fy != nullptr) + (s.fz != nullptr);
  FC_REQUIRE(faces == 0 || faces == 3, "gcr_step: face coefficients must be all set or all null");
  FC_REQUIRE(s.mask && (faces == 3 || (s.aE && s.aW && s.aN && s.aS && s.aU && s.aD && s.diagL)),
             "gcr_step: null coefficient");
  const int64_t n = (int64_t)s.nx * s.ny * s.nz;
  FC_REQUIRE(n + (int64_t)kMaxBlocks * kBlock < ((int64_t)1 << 31), "gcr_step: %lld cells exceed the 32-bit index range",
             (long long)n);
  Dirs d{};
  d.m = m;
  for (int j = 0; j < m; ++j) {
    FC_REQUIRE(Wd[j] && Pd[j], "gcr_step: null direction %d", j);
    FC_REQUIRE(ww_slot[j] >= K_WWHIST0 && ww_slot[j] < K_SLOTS, "gcr_step: history slot out of range");
    d.W[j] = Wd[j];
    d.P[j] = Pd[j];
    d.slot[j] = ww_slot[j];
  }
  const unsigned g = grid_of(n);
  hipLaunchKernelGGL(faces ? apply_kernel<true> : apply_kernel<false>, dim3(grid_for(n, kBlock)), dim3(kBlock), 0,
                     stream, z32, s, p, w);
  if (int rc = check_launch("gcr_step/apply")) return rc;
  if (m > 0) {
    hipLaunchKernelGGL(dots_kernel, dim3(g), dim3(kBlock), 0, stream, (const double*)w, d, n, partial);
    if (int rc = check_launch("gcr_step/dots")) return rc;
    hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const double*)partial, (int)g, kMaxDirs, m,
                       scal, (int)K_DOT0);
    if (int rc = check_launch("gcr_step/fold dots")) return rc;
  }
  hipLaunchKernelGGL(orth_kernel, dim3(g), dim3(kBlock), 0, stream, w, p, (const double*)r, d, (const double*)scal,
                     n, partial);
  if (int rc = check_launch("gcr_step/orth")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const double*)partial, (int)g, 2, 2, scal,
                     (int)K_RW);
  if (int rc = check_launch("gcr_step/fold orth")) return rc;
  hipLaunchKernelGGL(copy_scalar_kernel, dim3(1), dim3(1), 0, stream, scal, (int)K_WW, store_slot);
  if (int rc = check_launch("gcr_step/store")) return rc;
  hipLaunchKernelGGL(update_kernel, dim3(g), dim3(kBlock), 0, stream, x, r, r32, (const double*)p,
                     (const double*)w, (const double*)scal, n, partial);
  if (int rc = check_launch("gcr_step/update")) return rc;
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const double*)partial, (int)g, 1, 1, scal,
                     (int)K_RR);
  return check_launch("gcr_step/fold update");
}


int fc_fgmres_max_basis(void) { return kMaxBasis; }
int fc_fgmres_scalar_slots(void) { return F_SLOTS; }
int fc_fgmres_dot_slot(int i) { return F_DOT0 + i; }
int fc_fgmres_norm_slot(int i) { return F_NV0 + i; }
int fc_fgmres_partials_count(int64_t n) {
  const unsigned g = grid_for(n, kBlock);
  return (int)(g < kArnoldiBlocks ? g : kArnoldiBlocks) * (kMaxBasis + 1);
}

// One Arnoldi step of flexible GMRES: V[j+1] = A z32 orthogonalised against V[0..j]
// (classical Gram-Schmidt, coefficients w.V_i / V_i.V_i), v32 = (float) V[j+1];
// scal[dot_slot(i)] = w.V_i (w before the orthogonalisation), scal[norm_slot(j+1)] =
// ||V[j+1]||^2. scal[norm_slot(i)], i <= j, must hold ||V_i||^2.
int fc_fgmres_arnoldi(const float* z32, const fc_gcr_stencil* st, double* const* V, int j, float* v32,
                      double* partial, double* scal, hipStream_t stream) {
  FC_REQUIRE(z32 && st && V && v32 && partial && scal, "fgmres_arnoldi: null pointer argument");
  FC_REQUIRE(j >= 0 && j < kMaxBasis, "fgmres_arnoldi: 0 <= j < %d", kMaxBasis);
  FC_REQUIRE(st->nx >= 2 && st->ny >= 2 && st->nz >= 2, "fgmres_arnoldi: every extent must be >= 2");
  Stencil s{st->aE, st->aW, st->aN, st->aS, st->aU, st->aD, st->diagL, st->mask, st->nx, st->ny, st->nz,
            st->ax, st->out_idx, st->dt, st->fx, st->fy, st->fz};
  const int faces = (s.fx != nullptr) + (s.fy != nullptr) + (s.fz != nullptr);
  FC_REQUIRE(faces == 0 || faces == 3, "fgmres_arnoldi: face coefficients must be all set or all null");
  FC_REQUIRE(s.mask && (faces == 3 || (s.aE && s.aW && s.aN && s.aS && s.aU && s.aD && s.diagL)),
             "fgmres_arnoldi: null coefficient");
  const int64_t n = (int64_t)s.nx * s.ny * s.nz;
  FC_REQUIRE(n + (int64_t)kArnoldiBlocks * kBlock < ((int64_t)1 << 31),
             "fgmres_arnoldi: %lld cells exceed the 32-bit index range", (long long)n);
  Basis bs{};
  for (int i = 0; i <= j + 1; ++i) {
    FC_REQUIRE(V[i], "fgmres_arnoldi: null basis vector %d", i);
    bs.V[i] = V[i];
  }
  const unsigned g0 = grid_for(n, kBlock), g = g0 < kArnoldiBlocks ? g0 : kArnoldiBlocks;
  double* w = V[j + 1];
  ArnoldiLaunch<kMaxBasis>::apply(j + 1, faces == 3, g, stream, z32, s, w, bs, partial);
  if (int rc = check_launch("fgmres_arnoldi/apply")) return rc;
  hipLaunchKernelGGL(fold_columns_kernel, dim3(j + 1), dim3(kBlock), 0, stream, (const double*)partial, (int)g, scal,
                     (int)F_DOT0);
  if (int rc = check_launch("fgmres_arnoldi/fold dots")) return rc;
  ArnoldiLaunch<kMaxBasis>::orth(j + 1, g, stream, w, v32, bs, (const double*)scal, n, partial);
  if (int rc = check_launch("fgmres_arnoldi/orth")) return rc;
  hipLaunchKernelGGL(fold_columns_kernel, dim3(1), dim3(kBlock), 0, stream, (const double*)partial, (int)g, scal,
                     (int)(F_NV0 + j + 1));
  return check_launch("fgmres_arnoldi/fold norm");
}

// x += sum_{i<k} a[i] Z[i] (a on the host, Z FP32 device vectors)
int fc_fgmres_update_x(double* x, const float* const* Z, const double* a, int k, int64_t n, hipStream_t stream) {
  FC_REQUIRE(x && Z && a, "fgmres_update_x: null pointer argument");
  FC_REQUIRE(k >= 0 && k <= kMaxBasis, "fgmres_update_x: 0 <= k <= %d", kMaxBasis);
  if (k == 0 || n == 0) return 0;
  ZSet zs{};
  for (int i = 0; i < k; ++i) {
    FC_REQUIRE(Z[i], "fgmres_update_x: null vector %d", i);
    zs.Z[i] = Z[i];
    zs.a[i] = a[i];
  }
  ArnoldiLaunch<kMaxBasis>::update(k, stream, x, zs, n);
  return check_launch("fgmres_update_x");
}


int fc_fgmres_givens_slots(void) { return G_SLOTS; }
int fc_fgmres_givens_layout(int* r, int* g, int* est, int* ld) {
  FC_REQUIRE(r && g && est && ld, "fgmres_givens_layout: null pointer argument");
  *r = G_R;
  *g = G_G;
  *est = G_EST;
  *ld = kMaxBasis + 1;
  return 0;
}

// Givens update of Arnoldi step j on the device (one thread): reads the scalars that
// fc_fgmres_arnoldi left in scal, updates state (fc_fgmres_givens_slots() doubles)
// and writes the residual estimate |g_{j+1}| / bnorm at the est slot.
int fc_fgmres_givens(const double* scal, double* state, int j, double bnorm, hipStream_t stream) {
  FC_REQUIRE(scal && state, "fgmres_givens: null pointer argument");
  FC_REQUIRE(j >= 0 && j < kMaxBasis, "fgmres_givens: 0 <= j < %d", kMaxBasis);
  FC_REQUIRE(bnorm > 0.0, "fgmres_givens: bnorm must be positive");
  hipLaunchKernelGGL(arnoldi_givens_kernel, dim3(1), dim3(64), 0, stream, scal, state, j, bnorm);
  return check_launch("fgmres_givens");
}


// ---- BiCGSTAB building blocks ----
int fc_krylov_apply(const float* z32, const fc_gcr_stencil* st, double* out, int use_mask, hipStream_t stream) {
  FC_REQUIRE(z32 && st && out, "krylov_apply: null pointer argument");
  FC_REQUIRE(st->nx >= 2 && st->ny >= 2 && st->nz >= 2, "krylov_apply: every extent must be >= 2");
  Stencil s{st->aE, st->aW, st->aN, st->aS, st->aU, st->aD, st->diagL, st->mask, st->nx, st->ny, st->nz,
            st->ax, st->out_idx, st->dt, st->fx, st->fy, st->fz};
  const int faces = (s.fx != nullptr) + (s.fy != nullptr) + (s.fz != nullptr);
  FC_REQUIRE(faces == 0 || faces == 3, "krylov_apply: face coefficients must be all set or all null");
  FC_REQUIRE((!use_mask || s.mask) && (faces == 3 || (s.aE && s.aW && s.aN && s.aS && s.aU && s.aD && s.diagL)),
             "krylov_apply: null coefficient");
  const int64_t n = (int64_t)s.nx * s.ny * s.nz;
  FC_REQUIRE(n < ((int64_t)1 << 31), "krylov_apply: %lld cells exceed the 32-bit index range", (long long)n);
  hipLaunchKernelGGL(faces ? krylov_apply_kernel<true> : krylov_apply_kernel<false>, dim3(grid_for(n, kBlock)),
                     dim3(kBlock), 0, stream, z32, s, out, use_mask);
  return check_launch("krylov_apply");
}

// scal[slot0 + i] = a[i] . b[i] for i < npairs (1..3), FP64, fixed summation order;
// partial holds fc_gcr_partials_count(n) doubles
int fc_krylov_dots(const double* const* a, const double* const* b, int npairs, int64_t n, double* partial,
                   double* scal, int slot0, hipStream_t stream) {
  FC_REQUIRE(a && b && partial && scal, "krylov_dots: null pointer argument");
  FC_REQUIRE(npairs >= 1 && npairs <= 3, "krylov_dots: 1 <= npairs <= 3");
  FC_REQUIRE(slot0 >= 0 && slot0 + npairs <= K_SLOTS, "krylov_dots: slots out of range");
  DotPairs d{};
  for (int q = 0; q < npairs; ++q) {
    FC_REQUIRE(a[q] && b[q], "krylov_dots: null vector in pair %d", q);
    d.a[q] = a[q];
    d.b[q] = b[q];
  }
  const unsigned g = grid_of(n);
  if (npairs == 1)
    hipLaunchKernelGGL(krylov_dots_kernel<1>, dim3(g), dim3(kBlock), 0, stream, d, n, partial);
  else if (npairs == 2)
    hipLaunchKernelGGL(krylov_dots_kernel<2>, dim3(g), dim3(kBlock), 0, stream, d, n, partial);
  else
    hipLaunchKernelGGL(krylov_dots_kernel<3>, dim3(g), dim3(kBlock), 0, stream, d, n, partial);
  if (int rc = check_launch("krylov_dots")) return rc;
  hipLaunchKernelGGL(fold_columns_kernel, dim3(npairs), dim3(kBlock), 0, stream, (const double*)partial, (int)g, scal,
                     slot0);
  return check_launch("krylov_dots/fold");
}

int fc_bicgstab_direction(double* p, const double* r, const double* v, const float* mask, double beta, double omega,
                          float* p32, int64_t n, hipStream_t stream) {
  FC_REQUIRE(p && r && v && mask && p32, "bicgstab_direction: null pointer argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(bicgstab_dir_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, p, r, v, mask, beta,
                     beta * omega, p32, n);
  return check_launch("bicgstab_direction");
}

int fc_bicgstab_s(double* s, const double* r, const double* v, const float* mask, double alpha, float* s32, int64_t n,
                  hipStream_t stream) {
  FC_REQUIRE(s && r && v && mask && s32, "bicgstab_s: null pointer argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(bicgstab_s_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, s, r, v, mask, alpha,
                     s32, n);
  return check_launch("bicgstab_s");
}

int fc_bicgstab_update(double* x, double* r, const double* s, const double* t, const float* p_hat,
                       const float* s_hat, const float* mask, double alpha, double omega, int64_t n,
                       hipStream_t stream) {
  FC_REQUIRE(x && p_hat && mask, "bicgstab_update: null pointer argument");
  FC_REQUIRE(!s_hat || (r && s && t), "bicgstab_update: r, s and t are needed with s_hat");
  if (n == 0) return 0;
  hipLaunchKernelGGL(bicgstab_x_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x, r, s, t, p_hat, s_hat,
                     mask, alpha, omega, n);
  return check_launch("bicgstab_update");
}

}  // extern "C"
