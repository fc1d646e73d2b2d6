// thermal.hip — element-wise steps either side of the multigrid solve
// (behaviour of solver/thermal_kernels.cu:31-189):
//  * thermal_rhs:         rhs = m ((T_old + dt src) - LT0), theta0 = m (T_old - T0)
//  * thermal_reconstruct: T_new = T0 + theta, Neumann faces copy their interior
//                         neighbour (x, then y, then z face priority)
//  * convergence_metrics: max |T_new - T_old| over solid and over fluid nodes
//  * physics_check:       flag = any(!finite(T) || T < T_min || T > T_max)
// Reductions are two-stage (wavefront max/any, then one fold block), so the
// results do not depend on scheduling.
#include "fc_common.hpp"

using namespace fc;

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void rhs_kernel(const float* __restrict__ T_old, const float* __restrict__ T0,
                                                     const float* __restrict__ src, const float* __restrict__ LT0,
                                                     const float* __restrict__ mask, float* __restrict__ rhs,
                                                     float* __restrict__ theta0, float dt, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float m = mask[i], to = T_old[i];
  rhs[i] = m * ((to + dt * src[i]) - LT0[i]);
  theta0[i] = m * (to - T0[i]);
}

__global__ __launch_bounds__(kBlock) void reconstruct_kernel(const float* __restrict__ T0,
                                                             const float* __restrict__ theta,
                                                             float* __restrict__ T_new, int nx, int ny, int nz, int ax,
                                                             int out_idx) {
  const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t nynz = (int64_t)ny * nz;
  if (idx >= (int64_t)nx * nynz) return;
  const int i = (int)(idx / nynz), j = (int)((idx / nz) % ny), k = (int)(idx % nz);
  int64_t src = idx;
  bool neu = false;
  // a low face is Neumann unless it is the inlet of the flow axis
  auto low_neumann = [&](int this_ax) { return this_ax != ax || out_idx == 0; };
  if (i == nx - 1) { neu = true; src = (int64_t)(i - 1) * nynz + (int64_t)j * nz + k; }
  else if (i == 0 && low_neumann(0)) { neu = true; src = nynz + (int64_t)j * nz + k; }
  if (!neu) {
    if (j == ny - 1) { neu = true; src = (int64_t)i * nynz + (int64_t)(j - 1) * nz + k; }
    else if (j == 0 && low_neumann(1)) { neu = true; src = (int64_t)i * nynz + nz + k; }
  }
  if (!neu) {
    if (k == nz - 1) { neu = true; src = (int64_t)i * nynz + (int64_t)j * nz + (k - 1); }
    else if (k == 0 && low_neumann(2)) { neu = true; src = (int64_t)i * nynz + (int64_t)j * nz + 1; }
  }
  T_new[idx] = T0[src] + theta[src];
}

__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__global__ __launch_bounds__(kBlock) void dT_partial_kernel(const float* __restrict__ Tn, const float* __restrict__ To,
                                                            const uint8_t* __restrict__ solid, int64_t n,
                                                            float* __restrict__ partial) {
  __shared__ float sh[2][kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  float ms = 0.f, mf = 0.f;
  if (i < n) {
    const float d = fabsf(Tn[i] - To[i]);
    if (solid[i]) ms = d; else mf = d;
  }
  ms = wave_max(ms);
  mf = wave_max(mf);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = ms; sh[1][threadIdx.x >> 6] = mf; }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.f, b = 0.f;
    for (int w = 0; w < kBlock / 64; ++w) { a = fmaxf(a, sh[0][w]); b = fmaxf(b, sh[1][w]); }
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

__global__ __launch_bounds__(kBlock) void max_fold_kernel(const float* __restrict__ partial, int nparts,
                                                          float* __restrict__ out_solid, float* __restrict__ out_fluid) {
  __shared__ float sh[2][kBlock / 64];
  float a = 0.f, b = 0.f;
  for (int p = threadIdx.x; p < nparts; p += kBlock) {
    a = fmaxf(a, partial[2 * p]);
    b = fmaxf(b, partial[2 * p + 1]);
  }
  a = wave_max(a);
  b = wave_max(b);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = a; sh[1][threadIdx.x >> 6] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    float x = 0.f, y = 0.f;
    for (int w = 0; w < kBlock / 64; ++w) { x = fmaxf(x, sh[0][w]); y = fmaxf(y, sh[1][w]); }
    out_solid[0] = x;
    out_fluid[0] = y;
  }
}

__global__ __launch_bounds__(kBlock) void physics_check_kernel(const float* __restrict__ T, float T_min, float T_max,
                                                               int64_t n, int* __restrict__ partial) {
  __shared__ int sh[kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  bool bad = false;
  if (i < n) {
    const float v = T[i];
    bad = !isfinite(v) || v < T_min || v > T_max;
  }
  const uint64_t any = __ballot(bad);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = any != 0ull;
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0;
    for (int w = 0; w < kBlock / 64; ++w) a |= sh[w];
    partial[blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(kBlock) void any_fold_kernel(const int* __restrict__ partial, int nparts,
                                                          int* __restrict__ flag) {
  __shared__ int sh[kBlock / 64];
  int a = 0;
  for (int p = threadIdx.x; p < nparts; p += kBlock) a |= partial[p];
  const uint64_t any = __ballot(a != 0);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = any != 0ull;
  __syncthreads();
  if (threadIdx.x == 0) {
    int r = 0;
    for (int w = 0; w < kBlock / 64; ++w) r |= sh[w];
    flag[0] = r;
  }
}

}  // namespace

extern "C" {

int fc_thermal_rhs(const float* T_old, const float* T0, const float* src, const float* LT0, const float* mask,
                   float* rhs, float* theta0, double dt, int64_t n, hipStream_t stream) {
  FC_REQUIRE(n >= 0, "thermal_rhs: n must be >= 0");
  if (n == 0) return 0;
  FC_REQUIRE(T_old && T0 && src && LT0 && mask && rhs && theta0, "thermal_rhs: null pointer argument");
  hipLaunchKernelGGL(rhs_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, T_old, T0, src, LT0, mask, rhs,
                     theta0, (float)dt, n);
  return check_launch("thermal_rhs");
}

int fc_thermal_reconstruct(const float* T0, const float* theta, float* T_new, int nx, int ny, int nz, int ax,
                           int out_idx, hipStream_t stream) {
  FC_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "thermal_reconstruct: extents must be >= 2");
  FC_REQUIRE(ax >= 0 && ax <= 2, "thermal_reconstruct: axis must be 0, 1 or 2");
  FC_REQUIRE(T0 && theta && T_new, "thermal_reconstruct: null pointer argument");
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(reconstruct_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, T0, theta, T_new, nx,
                     ny, nz, ax, out_idx);
  return check_launch("thermal_reconstruct");
}

size_t fc_thermal_metrics_workspace_bytes(int64_t n) { return sizeof(float) * 2 * (size_t)grid_for(n, kBlock); }

int fc_thermal_convergence_metrics(const float* T_new, const float* T_old, const uint8_t* solid, int64_t n,
                                   float* out_solid, float* out_fluid, void* workspace, size_t ws_bytes,
                                   hipStream_t stream) {
  FC_REQUIRE(n > 0, "convergence_metrics: n must be > 0");
  FC_REQUIRE(T_new && T_old && solid && out_solid && out_fluid, "convergence_metrics: null pointer argument");
  const size_t need = fc_thermal_metrics_workspace_bytes(n);
  FC_REQUIRE(workspace && ws_bytes >= need, "convergence_metrics: workspace of %zu bytes required", need);
  const unsigned g = grid_for(n, kBlock);
  hipLaunchKernelGGL(dT_partial_kernel, dim3(g), dim3(kBlock), 0, stream, T_new, T_old, solid, n, (float*)workspace);
  if (int rc = check_launch("convergence_metrics")) return rc;
  hipLaunchKernelGGL(max_fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const float*)workspace, (int)g, out_solid,
                     out_fluid);
  return check_launch("convergence_metrics/fold");
}

size_t fc_thermal_check_workspace_bytes(int64_t n) { return sizeof(int) * (size_t)grid_for(n, kBlock); }

int fc_thermal_physics_check(const float* T, double T_min, double T_max, int64_t n, int* flag, void* workspace,
                             size_t ws_bytes, hipStream_t stream) {
  FC_REQUIRE(n > 0, "physics_check: n must be > 0");
  FC_REQUIRE(T && flag, "physics_check: null pointer argument");
  const size_t need = fc_thermal_check_workspace_bytes(n);
  FC_REQUIRE(workspace && ws_bytes >= need, "physics_check: workspace of %zu bytes required", need);
  const unsigned g = grid_for(n, kBlock);
  hipLaunchKernelGGL(physics_check_kernel, dim3(g), dim3(kBlock), 0, stream, T, (float)T_min, (float)T_max, n,
                     (int*)workspace);
  if (int rc = check_launch("physics_check")) return rc;
  hipLaunchKernelGGL(any_fold_kernel, dim3(1), dim3(kBlock), 0, stream, (const int*)workspace, (int)g, flag);
  return check_launch("physics_check/fold");
}

}  // extern "C"
