// lbm_ref.hip — per-operation D3Q19 kernels with the semantics of the
// reference extension lbm_kernels (one entry point per pybind function of
// lbm_kernels.cu:1347-1386). They keep the reference's two-buffer data flow
// (f -> f_post -> f) and serve the unchanged Python call surface; the fused
// in-place step lives in lbm_aa.hip.
//
// One thread per lattice node (or per boundary link / plane node), 256-thread
// blocks (four 64-wide wavefronts). All offsets are 64-bit so grids above
// 2^31/19 nodes are safe.
#include "fc_common.hpp"

#include <mutex>

using namespace fc;

namespace {

constexpr int kBlock = 256;

// Solid node: populations at rest equilibrium, macros rho=1, u=0
// (lbm_kernels.cu:277-283).
__device__ __forceinline__ void write_solid(float* __restrict__ f_post, float* rho, float* u, float* v,
                                            float* w, int64_t idx, int64_t N) {
#pragma unroll
  for (int q = 0; q < 19; ++q) f_post[q * N + idx] = WQ(q);
  rho[idx] = 1.f; u[idx] = 0.f; v[idx] = 0.f; w[idx] = 0.f;
}

__device__ __forceinline__ void load19(const float* __restrict__ f, int64_t idx, int64_t N, float fi[19]) {
#pragma unroll
  for (int q = 0; q < 19; ++q) fi[q] = f[q * N + idx];
}

// Pre-collision macros with inv_rho = 1/(rho+1e-12) (lbm_kernels.cu:286-299).
__device__ __forceinline__ void macros(const float fi[19], float& rho, float& ux, float& uy, float& uz) {
  float jx, jy, jz;
  moments0(fi, rho, jx, jy, jz);
  const float inv = 1.f / (rho + 1e-12f);
  ux = jx * inv; uy = jy * inv; uz = jz * inv;
}

// Collision kinds: 0 BGK, 1 TRT, 2 MRT, 3 MRT + Smagorinsky.
template <int KIND>
__global__ __launch_bounds__(kBlock) void macro_collide_kernel(
    const float* __restrict__ f, float* __restrict__ f_post, float* __restrict__ rho_o,
    float* __restrict__ u_o, float* __restrict__ v_o, float* __restrict__ w_o,
    const uint8_t* __restrict__ solid, float om1, float om2, const float* __restrict__ Sdev,
    float cs_dx2, int64_t N) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= N) return;
  if (solid[idx]) { write_solid(f_post, rho_o, u_o, v_o, w_o, idx, N); return; }
  float fi[19];
  load19(f, idx, N, fi);
  float rho, ux, uy, uz;
  macros(fi, rho, ux, uy, uz);
  rho_o[idx] = rho; u_o[idx] = ux; v_o[idx] = uy; w_o[idx] = uz;
  float feq[19];
  feq19(rho, ux, uy, uz, feq);

  if constexpr (KIND == 0) {  // BGK (lbm_kernels.cu:305-307)
#pragma unroll
    for (int q = 0; q < 19; ++q) f_post[q * N + idx] = fi[q] + om1 * (feq[q] - fi[q]);
  } else if constexpr (KIND == 1) {  // TRT (lbm_kernels.cu:355-364)
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const int o = OPP(q);
      const float fp = 0.5f * (fi[q] + fi[o]), fm = 0.5f * (fi[q] - fi[o]);
      const float ep = 0.5f * (feq[q] + feq[o]), em = 0.5f * (feq[q] - feq[o]);
      f_post[q * N + idx] = (fp + om1 * (ep - fp)) + (fm + om2 * (em - fm));
    }
  } else {  // MRT (lbm_kernels.cu:415-435), optional Smagorinsky (:1061-1116)
    float S[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) S[k] = Sdev[k];
    float fneq[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) fneq[q] = fi[q] - feq[q];
    if constexpr (KIND == 3) {
      // Frobenius norm of Pi_neq = sum_q fneq c c, then the Smagorinsky
      // quadratic for nu_eff; om1 is the base viscous rate omega0.
      float pxx = 0.f, pyy = 0.f, pzz = 0.f, pxy = 0.f, pxz = 0.f, pyz = 0.f;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const float x = (float)CX(q), y = (float)CY(q), z = (float)CZ(q);
        pxx += fneq[q] * x * x; pyy += fneq[q] * y * y; pzz += fneq[q] * z * z;
        pxy += fneq[q] * x * y; pxz += fneq[q] * x * z; pyz += fneq[q] * y * z;
      }
      const float pif = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.f * (pxy * pxy + pxz * pxz + pyz * pyz));
      const float inv_rho = 1.f / (rho + 1e-12f);
      const float nu0 = (1.f / om1 - 0.5f) / 3.f;
      const float disc = nu0 * nu0 + 4.f * cs_dx2 * (pif * inv_rho) * 4.5f;
      const float nu_eff = 0.5f * (nu0 + sqrtf(fmaxf(disc, nu0 * nu0)));
      const float om_eff = fminf(1.99f, 1.f / (3.f * nu_eff + 0.5f));
      S[9] = om_eff; S[11] = om_eff; S[13] = om_eff; S[14] = om_eff; S[15] = om_eff;
    }
    float df[19];
    mrt_df(fneq, S, df);
#pragma unroll
    for (int q = 0; q < 19; ++q) f_post[q * N + idx] = fi[q] - df[q];
  }
}

// Pull streaming with per-axis periodic wrap; a non-periodic out-of-range
// source keeps the node's own value (lbm_kernels.cu:447-496).
__global__ __launch_bounds__(kBlock) void stream_pull_kernel(const float* __restrict__ fin, float* __restrict__ fout,
                                                             int nx, int ny, int nz, int px, int py, int pz) {
  const int64_t N = (int64_t)nx * ny * nz;
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= N) return;
  const int64_t nynz = (int64_t)ny * nz;
  const int i = (int)(idx / nynz);
  const int j = (int)((idx / nz) % ny);
  const int k = (int)(idx % nz);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    int si = i - CX(q), sj = j - CY(q), sk = k - CZ(q);
    bool own = false;
    if (px) si = (si + nx) % nx; else own |= (si < 0 || si >= nx);
    if (py) sj = (sj + ny) % ny; else own |= (sj < 0 || sj >= ny);
    if (pz) sk = (sk + nz) % nz; else own |= (sk < 0 || sk >= nz);
    const int64_t src = own ? idx : ((int64_t)si * ny + sj) * nz + sk;
    fout[q * N + idx] = fin[q * N + src];
  }
}

// f <- f - sigma (f - feq(rho,u)) where sigma > 0 (lbm_kernels.cu:505-528).
__global__ __launch_bounds__(kBlock) void sponge_kernel(float* __restrict__ f, const float* __restrict__ rho,
                                                        const float* __restrict__ u, const float* __restrict__ v,
                                                        const float* __restrict__ w, const float* __restrict__ sigma,
                                                        int64_t N) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= N) return;
  const float sg = sigma[idx];
  if (sg <= 0.f) return;
  float feq[19];
  feq19(rho[idx], u[idx], v[idx], w[idx], feq);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const float fv = f[q * N + idx];
    f[q * N + idx] = fv - sg * (fv - feq[q]);
  }
}

// Bouzidi per link, writing population dir_i at the fluid node from the
// post-collision buffer (lbm_kernels.cu:542-588 and the direction choice of
// LBM_CHT.py:1814-1840).
__global__ __launch_bounds__(kBlock) void ibb_kernel(const float* __restrict__ f_post, float* __restrict__ f_stream,
                                                     const int32_t* __restrict__ fluid_flat,
                                                     const int32_t* __restrict__ dir_i,
                                                     const int32_t* __restrict__ dir_opp,
                                                     const float* __restrict__ w1, const float* __restrict__ w2,
                                                     const uint8_t* __restrict__ q_ge_half,
                                                     const int32_t* __restrict__ ff_flat,
                                                     const uint8_t* __restrict__ ff_valid, int64_t n_links,
                                                     int64_t N) {
  const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (l >= n_links) return;
  const int64_t xf = fluid_flat[l];
  const int qi = dir_i[l], qo = dir_opp[l];
  const float a = f_post[(int64_t)qo * N + xf];
  float val;
  if (q_ge_half[l]) val = w1[l] * a + w2[l] * f_post[(int64_t)qi * N + xf];
  else if (ff_valid[l]) val = w1[l] * a + w2[l] * f_post[(int64_t)qo * N + ff_flat[l]];
  else val = a;
  f_stream[(int64_t)qi * N + xf] = val;
}

// Solid nodes back to rest equilibrium (lbm_kernels.cu:597-609).
__global__ __launch_bounds__(kBlock) void reset_solid_kernel(float* __restrict__ f, const uint8_t* __restrict__ solid,
                                                             int64_t N) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= N || !solid[idx]) return;
#pragma unroll
  for (int q = 0; q < 19; ++q) f[q * N + idx] = WQ(q);
}

__device__ __forceinline__ void plane_nodes(int axis, int plane, int interior, int pa, int pb, int ny, int nz,
                                            int64_t& fb, int64_t& fi) {
  const int64_t nynz = (int64_t)ny * nz;
  if (axis == 0) { fb = plane * nynz + (int64_t)pa * nz + pb; fi = interior * nynz + (int64_t)pa * nz + pb; }
  else if (axis == 1) { fb = pa * nynz + (int64_t)plane * nz + pb; fi = pa * nynz + (int64_t)interior * nz + pb; }
  else { fb = pa * nynz + (int64_t)pb * nz + plane; fi = pa * nynz + (int64_t)pb * nz + interior; }
}

// Second-order Hermite tensor Q_ab = c_a c_b - delta_ab/3 with the
// reference's literals (lbm_kernels.cu:836-841).
__device__ __forceinline__ float qdiag(int c) { return c != 0 ? 0.66666667f : -0.33333333f; }

// Regularised Zou/He on one plane (behaviour of lbm_kernels.cu:791-969):
// bc_type 0 velocity inlet (rho from the known populations, clamped to
// [0.2, 5]); bc_type 1 pressure outlet (normal velocity from the knowns,
// backflow copies the interior node, tangential velocity from the interior
// node's moments). Unknown populations by Zou/He bounce of non-equilibrium,
// then the whole node is rebuilt from feq + regularised Pi_neq.
__global__ __launch_bounds__(kBlock) void zou_he_kernel(float* __restrict__ f, const uint8_t* __restrict__ solid,
                                                        int nx, int ny, int nz, int axis, int plane, int interior,
                                                        int bc_type, float u_in, float rho_out, int sign, int n1,
                                                        int n2) {
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  if (tid >= n1 * n2) return;
  const int64_t N = (int64_t)nx * ny * nz;
  int64_t fb, fin;
  plane_nodes(axis, plane, interior, tid / n2, tid % n2, ny, nz, fb, fin);
  if (solid[fb]) return;
  float fi[19];
  load19(f, fb, N, fi);
  const int unknown = plane == 0 ? 1 : -1;
  const int known = -unknown;
  float s0 = 0.f, sk = 0.f;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const int c = C_AX(q, axis);
    if (c == 0) s0 += fi[q];
    if (c == known) sk += fi[q];
  }
  float rho_bc, uvec[3] = {0.f, 0.f, 0.f};
  if (bc_type == 0) {
    rho_bc = (s0 + 2.f * sk) / (1.f - fabsf(u_in) + 1e-12f);
    rho_bc = fminf(fmaxf(rho_bc, 0.2f), 5.0f);
    uvec[axis] = u_in;
  } else {
    rho_bc = rho_out;
    const float uax = (float)known * ((s0 + 2.f * sk) / (rho_bc + 1e-12f) - 1.f);
    if (uax * (float)sign < 0.f) {
#pragma unroll
      for (int q = 0; q < 19; ++q) f[q * N + fb] = f[q * N + fin];
      return;
    }
    float gi[19];
    load19(f, fin, N, gi);
    float r, jx, jy, jz;
    moments0(gi, r, jx, jy, jz);
    const float inv = 1.f / (r + 1e-12f);
    uvec[0] = jx * inv; uvec[1] = jy * inv; uvec[2] = jz * inv;
    uvec[axis] = uax;
  }
  float feq[19];
  feq19(rho_bc, uvec[0], uvec[1], uvec[2], feq);
#pragma unroll
  for (int q = 0; q < 19; ++q)
    if (C_AX(q, axis) == unknown) fi[q] = fi[OPP(q)] + feq[q] - feq[OPP(q)];
  float pxx = 0.f, pyy = 0.f, pzz = 0.f, pxy = 0.f, pxz = 0.f, pyz = 0.f;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const float fn = fi[q] - feq[q];
    pxx += fn * qdiag(CX(q)); pyy += fn * qdiag(CY(q)); pzz += fn * qdiag(CZ(q));
    pxy += fn * (float)(CX(q) * CY(q)); pxz += fn * (float)(CX(q) * CZ(q)); pyz += fn * (float)(CY(q) * CZ(q));
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const float s = qdiag(CX(q)) * pxx + qdiag(CY(q)) * pyy + qdiag(CZ(q)) * pzz +
                    2.f * ((float)(CX(q) * CY(q)) * pxy + (float)(CX(q) * CZ(q)) * pxz + (float)(CY(q) * CZ(q)) * pyz);
    f[q * N + fb] = feq[q] + WQ(q) * s * 4.5f;
  }
}

// Convective outlet, pass 1: zero-gradient copy of the unknown populations
// from the interior plane, then per-block partial sums of the outlet density
// and fluid-node count (lbm_kernels.cu:1170-1238). Partials are FP64 and are
// combined in a fixed order in pass 2, so the correction is deterministic.
__global__ __launch_bounds__(kBlock) void convective_pass1(float* __restrict__ f, const uint8_t* __restrict__ solid,
                                                           int nx, int ny, int nz, int axis, int outlet,
                                                           int interior, int unknown_cax, int n1, int n2,
                                                           double* __restrict__ partial) {
  __shared__ double s_rho[kBlock / 64];
  __shared__ double s_cnt[kBlock / 64];
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  const int64_t N = (int64_t)nx * ny * nz;
  double r = 0.0, c = 0.0;
  if (tid < n1 * n2) {
    int64_t fo, fi;
    plane_nodes(axis, outlet, interior, tid / n2, tid % n2, ny, nz, fo, fi);
    if (!solid[fo]) {
      float rho = 0.f;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        float v;
        if (C_AX(q, axis) == unknown_cax) { v = f[q * N + fi]; f[q * N + fo] = v; }
        else v = f[q * N + fo];
        rho += v;
      }
      r = rho; c = 1.0;
    }
  }
  // wavefront reduction with cross-lane shuffles, then across the 4 waves
  for (int off = 32; off > 0; off >>= 1) {
    r += __shfl_down(r, off, 64);
    c += __shfl_down(c, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s_rho[wid] = r; s_cnt[wid] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) { a += s_rho[i]; b += s_cnt[i]; }
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

// Convective outlet, pass 2: every block folds the partials in the same order,
// then scales the outlet populations by rho_out * count / sum
// (lbm_kernels.cu:1249-1283, same degenerate-case guard).
__global__ __launch_bounds__(kBlock) void convective_pass2(float* __restrict__ f, const uint8_t* __restrict__ solid,
                                                           int nx, int ny, int nz, int axis, int outlet, int n1,
                                                           int n2, float rho_out, const double* __restrict__ partial,
                                                           int nparts) {
  __shared__ float s_corr;
  __shared__ int s_skip;
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nparts; ++i) { a += partial[2 * i]; b += partial[2 * i + 1]; }
    const float cnt = (float)b, sum = (float)a;
    s_skip = (cnt < 0.5f || sum < 1e-10f);
    s_corr = s_skip ? 1.f : rho_out * cnt / sum;
  }
  __syncthreads();
  if (s_skip) return;
  const float corr = s_corr;
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  if (tid >= n1 * n2) return;
  const int64_t N = (int64_t)nx * ny * nz;
  int64_t fo, fi;
  plane_nodes(axis, outlet, outlet, tid / n2, tid % n2, ny, nz, fo, fi);
  if (solid[fo]) return;
#pragma unroll
  for (int q = 0; q < 19; ++q) f[q * N + fo] *= corr;
}

// Scratch of fc_lbm_convective_bc_noalloc: the FP64 partial sums (outlet density,
// fluid-node count) of convective_pass1, two per block. Planes of up to
// kConvScratchBlocks * kBlock = 1,048,576 nodes; 64 KiB held by every device the
// library runs on until the process ends. Pass 1 writes every slot that pass 2
// reads, so nothing has to start at zero.
constexpr int kConvScratchBlocks = 4096;
static __device__ __align__(256) double scratch[2 * kConvScratchBlocks];

// Device address of `scratch` per device (the host-side name is only a stand-in).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int convective_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(2, "convective_bc_noalloc: hipGetDevice failed: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_error(1, "convective_bc_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(2, "convective_bc_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// Argument checks shared by both convective entry points.
int convective_args(int nx, int ny, int nz, int axis, int outlet_idx, int interior_idx, int unknown_cax, int n1,
                    int n2) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "convective_bc: grid extents must be positive");
  FC_REQUIRE(axis >= 0 && axis <= 2, "convective_bc: axis must be 0, 1 or 2");
  FC_REQUIRE(unknown_cax == 1 || unknown_cax == -1, "convective_bc: unknown_cax must be +1 or -1");
  const int e1 = axis == 0 ? ny : nx, e2 = axis == 2 ? ny : nz;
  FC_REQUIRE(n1 == e1 && n2 == e2, "convective_bc: plane dims (%d,%d) do not match axis %d (%d,%d)", n1, n2, axis, e1,
             e2);
  const int n_ax = axis == 0 ? nx : (axis == 1 ? ny : nz);
  FC_REQUIRE(outlet_idx >= 0 && outlet_idx < n_ax && interior_idx >= 0 && interior_idx < n_ax,
             "convective_bc: plane/interior index out of range");
  return 0;
}

int convective_launch(float* f, const uint8_t* solid, int nx, int ny, int nz, int axis, int outlet_idx,
                      int interior_idx, int unknown_cax, float rho_out_lat, int n1, int n2, double* partial,
                      hipStream_t stream) {
  const unsigned g = grid_for((int64_t)n1 * n2, kBlock);
  hipLaunchKernelGGL(convective_pass1, dim3(g), dim3(kBlock), 0, stream, f, solid, nx, ny, nz, axis, outlet_idx,
                     interior_idx, unknown_cax, n1, n2, partial);
  int rc = check_launch("convective_bc/pass1");
  if (rc) return rc;
  hipLaunchKernelGGL(convective_pass2, dim3(g), dim3(kBlock), 0, stream, f, solid, nx, ny, nz, axis, outlet_idx, n1,
                     n2, rho_out_lat, (const double*)partial, (int)g);
  return check_launch("convective_bc/pass2");
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int fc_lbm_macro_collide_bgk(const float* f, float* f_post, float* rho, float* u, float* v, float* w,
                             const uint8_t* solid, float omega, int64_t N, hipStream_t stream) {
  FC_REQUIRE(N >= 0, "macro_collide_bgk: N must be >= 0");
  if (N == 0) return 0;
  FC_REQUIRE(f && f_post && rho && u && v && w && solid, "macro_collide_bgk: null pointer argument");
  hipLaunchKernelGGL(macro_collide_kernel<0>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f, f_post, rho,
                     u, v, w, solid, omega, 0.f, nullptr, 0.f, N);
  return check_launch("macro_collide_bgk");
}

int fc_lbm_macro_collide_trt(const float* f, float* f_post, float* rho, float* u, float* v, float* w,
                             const uint8_t* solid, float omega_plus, float omega_minus, int64_t N,
                             hipStream_t stream) {
  FC_REQUIRE(N >= 0, "macro_collide_trt: N must be >= 0");
  if (N == 0) return 0;
  FC_REQUIRE(f && f_post && rho && u && v && w && solid, "macro_collide_trt: null pointer argument");
  hipLaunchKernelGGL(macro_collide_kernel<1>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f, f_post, rho,
                     u, v, w, solid, omega_plus, omega_minus, nullptr, 0.f, N);
  return check_launch("macro_collide_trt");
}

int fc_lbm_macro_collide_mrt(const float* f, float* f_post, float* rho, float* u, float* v, float* w,
                             const uint8_t* solid, const float* S_rates, int64_t N, hipStream_t stream) {
  FC_REQUIRE(N >= 0, "macro_collide_mrt: N must be >= 0");
  if (N == 0) return 0;
  FC_REQUIRE(f && f_post && rho && u && v && w && solid && S_rates, "macro_collide_mrt: null pointer argument");
  hipLaunchKernelGGL(macro_collide_kernel<2>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f, f_post, rho,
                     u, v, w, solid, 0.f, 0.f, S_rates, 0.f, N);
  return check_launch("macro_collide_mrt");
}

int fc_lbm_macro_collide_mrt_smag(const float* f, float* f_post, float* rho, float* u, float* v, float* w,
                                  const uint8_t* solid, const float* S_rates, float Cs_dx2, float omega0, int64_t N,
                                  hipStream_t stream) {
  FC_REQUIRE(N >= 0, "macro_collide_mrt_smag: N must be >= 0");
  if (N == 0) return 0;
  FC_REQUIRE(f && f_post && rho && u && v && w && solid && S_rates, "macro_collide_mrt_smag: null pointer argument");
  FC_REQUIRE(omega0 > 0.f, "macro_collide_mrt_smag: omega0 must be > 0");
  hipLaunchKernelGGL(macro_collide_kernel<3>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f, f_post, rho,
                     u, v, w, solid, omega0, 0.f, S_rates, Cs_dx2, N);
  return check_launch("macro_collide_mrt_smag");
}

int fc_lbm_stream_pull(const float* f_in, float* f_out, int nx, int ny, int nz, int periodic_x, int periodic_y,
                       int periodic_z, hipStream_t stream) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "stream_pull: grid extents must be positive");
  FC_REQUIRE(f_in && f_out && f_in != f_out, "stream_pull: f_in and f_out must be distinct device buffers");
  const int64_t N = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(stream_pull_kernel, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f_in, f_out, nx, ny,
                     nz, periodic_x != 0, periodic_y != 0, periodic_z != 0);
  return check_launch("stream_pull");
}

int fc_lbm_sponge(float* f, const float* rho, const float* u, const float* v, const float* w, const float* sigma,
                  int64_t N, hipStream_t stream) {
  FC_REQUIRE(N >= 0, "sponge: N must be >= 0");
  if (N == 0) return 0;
  FC_REQUIRE(f && rho && u && v && w && sigma, "sponge: null pointer argument");
  hipLaunchKernelGGL(sponge_kernel, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f, rho, u, v, w, sigma, N);
  return check_launch("sponge");
}

int fc_lbm_ibb(const float* f_post, float* f_stream, const int32_t* fluid_flat, const int32_t* dir_i,
               const int32_t* dir_opp, const float* w1, const float* w2, const uint8_t* q_ge_half,
               const int32_t* ff_flat, const uint8_t* ff_valid, int64_t n_links, int64_t N, hipStream_t stream) {
  FC_REQUIRE(n_links >= 0 && N >= 0, "ibb: sizes must be >= 0");
  if (n_links == 0) return 0;  // lbm_kernels.cu:737
  FC_REQUIRE(f_post && f_stream && fluid_flat && dir_i && dir_opp && w1 && w2 && q_ge_half && ff_flat && ff_valid,
             "ibb: null pointer argument");
  hipLaunchKernelGGL(ibb_kernel, dim3(grid_for(n_links, kBlock)), dim3(kBlock), 0, stream, f_post, f_stream,
                     fluid_flat, dir_i, dir_opp, w1, w2, q_ge_half, ff_flat, ff_valid, n_links, N);
  return check_launch("ibb");
}

int fc_lbm_reset_solid(float* f, const uint8_t* solid, int64_t N, hipStream_t stream) {
  FC_REQUIRE(N >= 0, "reset_solid: N must be >= 0");
  if (N == 0) return 0;
  FC_REQUIRE(f && solid, "reset_solid: null pointer argument");
  hipLaunchKernelGGL(reset_solid_kernel, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, f, solid, N);
  return check_launch("reset_solid");
}

int fc_lbm_zou_he_bc(float* f, const uint8_t* solid, int nx, int ny, int nz, int axis, int plane_idx,
                     int interior_idx, int bc_type, float u_in_lat, float rho_out_lat, int sign,
                     hipStream_t stream) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "zou_he_bc: grid extents must be positive");
  FC_REQUIRE(axis >= 0 && axis <= 2, "zou_he_bc: axis must be 0, 1 or 2");
  FC_REQUIRE(bc_type == 0 || bc_type == 1, "zou_he_bc: bc_type must be 0 (velocity) or 1 (pressure)");
  const int n_ax = axis == 0 ? nx : (axis == 1 ? ny : nz);
  FC_REQUIRE(plane_idx >= 0 && plane_idx < n_ax && interior_idx >= 0 && interior_idx < n_ax,
             "zou_he_bc: plane/interior index out of range");
  FC_REQUIRE(f && solid, "zou_he_bc: null pointer argument");
  const int n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
  hipLaunchKernelGGL(zou_he_kernel, dim3(grid_for((int64_t)n1 * n2, kBlock)), dim3(kBlock), 0, stream, f, solid, nx,
                     ny, nz, axis, plane_idx, interior_idx, bc_type, u_in_lat, rho_out_lat, sign, n1, n2);
  return check_launch("zou_he_bc");
}

size_t fc_lbm_convective_bc_workspace_bytes(int n1, int n2) {
  if (n1 <= 0 || n2 <= 0) return 0;
  return sizeof(double) * 2 * (size_t)grid_for((int64_t)n1 * n2, kBlock);
}

int fc_lbm_convective_bc(float* f, const uint8_t* solid, int nx, int ny, int nz, int axis, int outlet_idx,
                         int interior_idx, int unknown_cax, float rho_out_lat, int n1, int n2, void* workspace,
                         size_t workspace_bytes, hipStream_t stream) {
  int rc = convective_args(nx, ny, nz, axis, outlet_idx, interior_idx, unknown_cax, n1, n2);
  if (rc) return rc;
  FC_REQUIRE(f && solid, "convective_bc: null pointer argument");
  const size_t need = fc_lbm_convective_bc_workspace_bytes(n1, n2);
  FC_REQUIRE(workspace && workspace_bytes >= need, "convective_bc: workspace of %zu bytes required", need);
  return convective_launch(f, solid, nx, ny, nz, axis, outlet_idx, interior_idx, unknown_cax, rho_out_lat, n1, n2,
                           (double*)workspace, stream);
}

int fc_lbm_convective_bc_noalloc(float* f, const uint8_t* solid, int nx, int ny, int nz, int axis, int outlet_idx,
                                 int interior_idx, int unknown_cax, float rho_out_lat, int n1, int n2,
                                 hipStream_t stream) {
  int rc = convective_args(nx, ny, nz, axis, outlet_idx, interior_idx, unknown_cax, n1, n2);
  if (rc) return rc;
  FC_REQUIRE((int64_t)n1 * n2 <= (int64_t)kConvScratchBlocks * kBlock,
             "convective_bc_noalloc: outlet plane of %lld nodes exceeds the %d-node scratch limit; "
             "use fc_lbm_convective_bc with a workspace", (long long)n1 * n2, kConvScratchBlocks * kBlock);
  FC_REQUIRE(f && solid, "convective_bc_noalloc: null pointer argument");
  double* partial = nullptr;
  rc = convective_scratch(&partial);
  if (rc) return rc;
  return convective_launch(f, solid, nx, ny, nz, axis, outlet_idx, interior_idx, unknown_cax, rho_out_lat, n1, n2,
                           partial, stream);
}

}  // extern "C"
