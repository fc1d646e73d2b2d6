// links.hip — Bouzidi link classification and wall distances on the GPU.
//
// Behaviour of pre_post/sdf_tracing.py:20-259:
//  * link_mask: for every fluid node, bit q (1..18) when the neighbour x + c_q
//    is solid (in range, or wrapped on axes flagged periodic — the reference
//    never wraps), and bit q of a second mask when x_ff = x - c_q is in range
//    (or wrapped) and fluid. The caller compacts the set bits per direction
//    in node order, which gives the reference's link order (q outer, C order
//    inner).
//  * link_q: coarse sampling of the trilinear SDF along the link, first
//    +/- crossing, bisection, q = hi / (L + 1e-30) (0.5 without a crossing),
//    and the BFL weights. Every float operation is rounded on its own
//    (contraction off), matching the eager FP32 operation sequence.
#include "fc_common.hpp"

#pragma clang fp contract(off)

using namespace fc;

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool nb(int c, int d, int n, int per, int& out) {
  int v = c + d;
  if (v < 0 || v >= n) {
    if (!per) return false;
    v = v < 0 ? v + n : v - n;
  }
  out = v;
  return true;
}

__global__ __launch_bounds__(kBlock) void link_mask_kernel(const uint8_t* __restrict__ solid, int nx, int ny, int nz,
                                                           int px, int py, int pz, uint32_t* __restrict__ mask,
                                                           uint32_t* __restrict__ ffok) {
  const int64_t N = (int64_t)nx * ny * nz;
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (x >= N) return;
  uint32_t m = 0, ok = 0;
  if (!solid[x]) {
    const int i = (int)(x / ((int64_t)ny * nz)), j = (int)((x / nz) % ny), k = (int)(x % nz);
#pragma unroll
    for (int q = 1; q < 19; ++q) {
      int a, b, c;
      if (nb(i, CX(q), nx, px, a) && nb(j, CY(q), ny, py, b) && nb(k, CZ(q), nz, pz, c) &&
          solid[((int64_t)a * ny + b) * nz + c]) {
        m |= 1u << q;
        int a2, b2, c2;
        if (nb(i, -CX(q), nx, px, a2) && nb(j, -CY(q), ny, py, b2) && nb(k, -CZ(q), nz, pz, c2) &&
            !solid[((int64_t)a2 * ny + b2) * nz + c2])
          ok |= 1u << q;
      }
    }
  }
  mask[x] = m;
  ffok[x] = ok;
}

struct Sdf {
  const float* phi;
  int nx, ny, nz;
  float ox, oy, oz, dx;
  float hx, hy, hz;  // extent - 1.001 (clamp limits)
  int px, py, pz;    // periodic axes wrap instead of clamping (link q search; build extension)
};

// grid coordinate along one axis, clamped to [0, n - 1.001] (the reference's rule)
__device__ __forceinline__ void axis_coord(float g, float h, int n, int& i0, int& i1, float& fr) {
  g = g < 0.0f ? 0.0f : (g > h ? h : g);
  i0 = (int)floorf(g);
  fr = g - (float)i0;
  i1 = min(i0 + 1, n - 1);
}

// trilinear SDF at a world point (sdf_trilinear, sdf_tracing.py:88-125)
__device__ float sdf_tri(const Sdf& s, float x, float y, float z) {
  const float gx = (x - s.ox) / s.dx, gy = (y - s.oy) / s.dx, gz = (z - s.oz) / s.dx;
  int i0, i1, j0, j1, k0, k1;
  float fx, fy, fz;
  axis_coord(gx, s.hx, s.nx, i0, i1, fx);
  axis_coord(gy, s.hy, s.ny, j0, j1, fy);
  axis_coord(gz, s.hz, s.nz, k0, k1, fz);
  auto P = [&](int a, int b, int c) { return s.phi[((int64_t)a * s.ny + b) * s.nz + c]; };
  const float c000 = P(i0, j0, k0), c100 = P(i1, j0, k0), c010 = P(i0, j1, k0), c110 = P(i1, j1, k0);
  const float c001 = P(i0, j0, k1), c101 = P(i1, j0, k1), c011 = P(i0, j1, k1), c111 = P(i1, j1, k1);
  const float cx00 = c000 + fx * (c100 - c000);
  const float cx10 = c010 + fx * (c110 - c010);
  const float cx01 = c001 + fx * (c101 - c001);
  const float cx11 = c011 + fx * (c111 - c011);
  const float cxy0 = cx00 + fy * (cx10 - cx00);
  const float cxy1 = cx01 + fy * (cx11 - cx01);
  return cxy0 + fz * (cxy1 - cxy0);
}

// Sample of the trilinear SDF at x0 + t d. Non-periodic axes follow the reference's
// arithmetic (world point, grid coordinate clamped to [0, n - 1.001]). Periodic
// axes (build extension) take the node's own index plus the offset t d / pitch,
// wrapped: the cell and the fraction do not depend on where the node sits, so
// every copy of a periodic geometry gets the same bits (a box of whole periods is
// an exact sub-domain of a larger box).
__device__ __forceinline__ void sample_axis(float x0, float m, float o, float pitch, float h, int n, int per, int node,
                                            int& i0, int& i1, float& fr) {
  if (per) {
    const float off = m / pitch;
    const float fl = floorf(off);
    fr = off - fl;
    int a = (node + (int)fl) % n;
    a = a < 0 ? a + n : a;
    i0 = a;
    i1 = a + 1 == n ? 0 : a + 1;
  } else {
    const float x = x0 + m;
    float g = (x - o) / pitch;
    g = g < 0.0f ? 0.0f : (g > h ? h : g);
    i0 = (int)floorf(g);
    fr = g - (float)i0;
    i1 = min(i0 + 1, n - 1);
  }
}

__device__ float sdf_sample(const Sdf& s, int ni, int nj, int nk, float x0, float y0, float z0, float t, float dx,
                            float dy, float dz) {
  int i0, i1, j0, j1, k0, k1;
  float fx, fy, fz;
  sample_axis(x0, t * dx, s.ox, s.dx, s.hx, s.nx, s.px, ni, i0, i1, fx);
  sample_axis(y0, t * dy, s.oy, s.dx, s.hy, s.ny, s.py, nj, j0, j1, fy);
  sample_axis(z0, t * dz, s.oz, s.dx, s.hz, s.nz, s.pz, nk, k0, k1, fz);
  auto P = [&](int a, int b, int c) { return s.phi[((int64_t)a * s.ny + b) * s.nz + c]; };
  const float c000 = P(i0, j0, k0), c100 = P(i1, j0, k0), c010 = P(i0, j1, k0), c110 = P(i1, j1, k0);
  const float c001 = P(i0, j0, k1), c101 = P(i1, j0, k1), c011 = P(i0, j1, k1), c111 = P(i1, j1, k1);
  const float cx00 = c000 + fx * (c100 - c000);
  const float cx10 = c010 + fx * (c110 - c010);
  const float cx01 = c001 + fx * (c101 - c001);
  const float cx11 = c011 + fx * (c111 - c011);
  const float cxy0 = cx00 + fy * (cx10 - cx00);
  const float cxy1 = cx01 + fy * (cx11 - cx01);
  return cxy0 + fz * (cxy1 - cxy0);
}

// q in (0, 1] along the segment x0 + t d, t in [0, L] (bfl_q_from_sdf,
// sdf_tracing.py:127-184): n_coarse + 1 samples at torch.linspace(0, 1) * L, the
// first + -> (not +) crossing, n_bisect halvings keeping the solid-side end hi,
// q = hi / (L + 1e-30); 0.5 without a crossing. (ni, nj, nk): the start node's
// index, read on periodic axes only.
__device__ float q_search(const Sdf& s, int ni, int nj, int nk, float x0, float y0, float z0, float dx, float dy,
                          float dz, float L, int n_coarse, int n_bisect) {
  // torch.linspace(0, 1, n_coarse + 1) in FP32
  const int steps = n_coarse + 1, half = steps / 2;
  const float step = 1.0f / (float)(steps - 1);
  float t_prev = 0.f, p_prev = 0.f, lo = 0.f, hi = 0.f;
  bool found = false;
  for (int k = 0; k < steps; ++k) {
    const float tv = k < half ? 0.0f + step * (float)k : 1.0f - step * (float)(steps - k - 1);
    const float t = tv * L;
    const float p = sdf_sample(s, ni, nj, nk, x0, y0, z0, t, dx, dy, dz);
    if (k > 0 && !found && p_prev > 0.0f && !(p > 0.0f)) {
      found = true;
      lo = t_prev;
      hi = t;
    }
    t_prev = t;
    p_prev = p;
  }
  if (!found) return 0.5f;
  for (int it = 0; it < n_bisect; ++it) {
    const float mid = 0.5f * (lo + hi);
    const float pm = sdf_sample(s, ni, nj, nk, x0, y0, z0, mid, dx, dy, dz);
    if (pm > 0.0f) lo = mid; else hi = mid;
  }
  return hi / (L + 1e-30f);
}

__device__ __forceinline__ void bfl_weights(float qv, float& w1, float& w2, uint8_t& ge) {
  const float a = qv >= 0.5f ? 1.0f / (2.0f * qv) : 2.0f * qv;
  w1 = a;
  w2 = 1.0f - a;
  ge = qv >= 0.5f;
}

__global__ __launch_bounds__(kBlock) void link_q_kernel(Sdf s, const int64_t* __restrict__ fluid_flat,
                                                        const int64_t* __restrict__ dir_i, int64_t n, int n_coarse,
                                                        int n_bisect, float* __restrict__ q_out,
                                                        float* __restrict__ w1, float* __restrict__ w2,
                                                        uint8_t* __restrict__ ge) {
  const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (l >= n) return;
  const int64_t f = fluid_flat[l];
  const int ni = (int)(f / ((int64_t)s.ny * s.nz)), nj = (int)((f / s.nz) % s.ny), nk = (int)(f % s.nz);
  const float fi = (float)ni, fj = (float)nj, fk = (float)nk;
  const float x0 = fi * s.dx + s.ox, y0 = fj * s.dx + s.oy, z0 = fk * s.dx + s.oz;
  const int q = (int)dir_i[l];
  const float cx = (float)CX(q), cy = (float)CY(q), cz = (float)CZ(q);
  const float cn = sqrtf(cx * cx + cy * cy + cz * cz);
  const float den = cn + 1e-30f;
  const float dx = cx / den, dy = cy / den, dz = cz / den;
  const float L = s.dx * cn;
  const float qv = q_search(s, ni, nj, nk, x0, y0, z0, dx, dy, dz, L, n_coarse, n_bisect);
  q_out[l] = qv;
  bfl_weights(qv, w1[l], w2[l], ge[l]);
}

// sdf_trilinear (sdf_tracing.py:88-125) at arbitrary points x[n][3]
__global__ __launch_bounds__(kBlock) void sdf_points_kernel(Sdf s, const float* __restrict__ x, int64_t n,
                                                            float* __restrict__ out) {
  const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (l >= n) return;
  out[l] = sdf_tri(s, x[3 * l], x[3 * l + 1], x[3 * l + 2]);
}

// bfl_q_from_sdf for arbitrary rays: start points x0[n][3], unit directions d[n][3], lengths L[n]
__global__ __launch_bounds__(kBlock) void q_rays_kernel(Sdf s, const float* __restrict__ x0,
                                                        const float* __restrict__ d, const float* __restrict__ L,
                                                        int64_t n, int n_coarse, int n_bisect,
                                                        float* __restrict__ q_out) {
  const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (l >= n) return;
  q_out[l] = q_search(s, 0, 0, 0, x0[3 * l], x0[3 * l + 1], x0[3 * l + 2], d[3 * l], d[3 * l + 1], d[3 * l + 2],
                      L[l], n_coarse, n_bisect);
}

__global__ __launch_bounds__(kBlock) void weights_kernel(const float* __restrict__ q, int64_t n,
                                                         uint8_t* __restrict__ ge, float* __restrict__ w1,
                                                         float* __restrict__ w2) {
  const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (l >= n) return;
  bfl_weights(q[l], w1[l], w2[l], ge[l]);
}

int make_sdf(Sdf& s, const float* phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch, int px,
             int py, int pz) {
  FC_REQUIRE(nx > 1 && ny > 1 && nz > 1, "sdf: phi extents must be > 1");
  FC_REQUIRE(pitch > 0.f, "sdf: pitch must be > 0");
  FC_REQUIRE(phi, "sdf: null phi");
  s.phi = phi; s.nx = nx; s.ny = ny; s.nz = nz; s.ox = ox; s.oy = oy; s.oz = oz; s.dx = pitch;
  s.hx = (float)((double)nx - 1.001); s.hy = (float)((double)ny - 1.001); s.hz = (float)((double)nz - 1.001);
  s.px = px != 0; s.py = py != 0; s.pz = pz != 0;
  return 0;
}

}  // namespace

extern "C" {

int fc_links_mask(const uint8_t* solid, int nx, int ny, int nz, int px, int py, int pz, uint32_t* mask,
                  uint32_t* ffok, hipStream_t stream) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "links_mask: grid extents must be positive");
  FC_REQUIRE(solid && mask && ffok, "links_mask: null pointer argument");
  const int64_t N = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(link_mask_kernel, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, solid, nx, ny, nz,
                     px != 0, py != 0, pz != 0, mask, ffok);
  return check_launch("links_mask");
}

int fc_links_q(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch, int px, int py,
               int pz, const int64_t* fluid_flat, const int64_t* dir_i, int64_t n, int n_coarse, int n_bisect,
               float* q, float* w1, float* w2, uint8_t* ge, hipStream_t stream) {
  FC_REQUIRE(nx > 1 && ny > 1 && nz > 1, "links_q: phi extents must be > 1");
  FC_REQUIRE(n_coarse >= 1 && n_bisect >= 0, "links_q: n_coarse >= 1 and n_bisect >= 0 required");
  FC_REQUIRE(pitch > 0.f, "links_q: pitch must be > 0");
  if (n == 0) return 0;
  FC_REQUIRE(phi && fluid_flat && dir_i && q && w1 && w2 && ge, "links_q: null pointer argument");
  Sdf s;
  if (int rc = make_sdf(s, phi, nx, ny, nz, ox, oy, oz, pitch, px, py, pz)) return rc;
  hipLaunchKernelGGL(link_q_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, s, fluid_flat, dir_i, n,
                     n_coarse, n_bisect, q, w1, w2, ge);
  return check_launch("links_q");
}

int fc_sdf_trilinear(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch,
                     const float* x, int64_t n, float* out, hipStream_t stream) {
  FC_REQUIRE(n >= 0, "sdf_trilinear: n must be >= 0");
  Sdf s;
  if (int rc = make_sdf(s, phi, nx, ny, nz, ox, oy, oz, pitch, 0, 0, 0)) return rc;
  if (n == 0) return 0;
  FC_REQUIRE(x && out, "sdf_trilinear: null pointer argument");
  hipLaunchKernelGGL(sdf_points_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, s, x, n, out);
  return check_launch("sdf_trilinear");
}

int fc_bfl_q_rays(const float* phi, int nx, int ny, int nz, float ox, float oy, float oz, float pitch,
                  const float* x0, const float* d, const float* L, int64_t n, int n_coarse, int n_bisect, float* q,
                  hipStream_t stream) {
  FC_REQUIRE(n >= 0 && n_coarse >= 1 && n_bisect >= 0, "bfl_q_rays: n >= 0, n_coarse >= 1, n_bisect >= 0 required");
  Sdf s;
  if (int rc = make_sdf(s, phi, nx, ny, nz, ox, oy, oz, pitch, 0, 0, 0)) return rc;
  if (n == 0) return 0;
  FC_REQUIRE(x0 && d && L && q, "bfl_q_rays: null pointer argument");
  hipLaunchKernelGGL(q_rays_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, s, x0, d, L, n, n_coarse,
                     n_bisect, q);
  return check_launch("bfl_q_rays");
}

int fc_bfl_weights(const float* q, int64_t n, uint8_t* ge, float* w1, float* w2, hipStream_t stream) {
  FC_REQUIRE(n >= 0, "bfl_weights: n must be >= 0");
  if (n == 0) return 0;
  FC_REQUIRE(q && ge && w1 && w2, "bfl_weights: null pointer argument");
  hipLaunchKernelGGL(weights_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, q, n, ge, w1, w2);
  return check_launch("bfl_weights");
}

}  // extern "C"
