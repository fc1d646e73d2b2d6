// d3q7.hip — D3Q7 advection-diffusion of the fluid temperature (conjugate heat
// transfer with the D3Q19 flow).
//
// Populations g_i, i = 0 (rest), 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z (the first seven
// D3Q19 directions); weights w_0 = 1/4, w_i = 1/8, so c_s^2 = 1/4 and the thermal
// diffusivity is kappa = (1/omega - 1/2) / 4. Equilibrium g_i^eq = w_i T (1 + 4 c_i.u)
// with the flow velocity u of the same step. Storage: post-collision values
// g*[7][N] in C order (z fastest), two buffers swapped by the caller.
//
// One launch per step, per node x (pull): input i is g*_i(x - c_i) of the previous
// step, except
//   * x - c_i is an adiabatic wall node: bounce-back, g*_opp(i)(x) (zero flux);
//   * x - c_i is a conducting solid node (kind 1): anti-bounce-back at its temperature,
//     -g*_opp(i)(x) + 2 w_i T_solid(x - c_i) (Dirichlet on the face between the nodes);
//   * x - c_i is outside the domain on a non-periodic axis: g*_i(x) (zero gradient);
// then T = sum_i g_i (written), g*_i = g_i + omega (g_i^eq - g_i) + w_i Q(x) with an
// optional heat source Q (temperature units per step). Nodes on the inlet plane
// (flag bit 4) are held at T_in: g*_i = w_i T_in (1 + 4 c_i.u). Kind 3 marks a
// conducting solid carried on the lattice (conjugate conduction without a separate
// solve): it streams and collides like fluid, with u = 0 and its own rate omega_s,
// so kappa_s = (1/omega_s - 1/2)/4; in steady state the face between two kinds
// conducts with the harmonic mean of the two sides, as a finite-volume operator
// with harmonic-mean face conductivities does. Wall nodes are not
// touched; a solid node writes T_out = T_solid and, when acc is given, adds to acc the
// energy its fluid neighbours (inlet nodes excepted) lose to it in this step: per link,
// g*_opp(i)(x + c_i) minus the anti-bounce-back value that replaces it, so the solid
// gains exactly what the fluid loses. Every sum and product rounds where it is
// written (no contraction).
#include "fc_common.hpp"

using namespace fc;

namespace {

constexpr int kBlock = 256;
constexpr uint8_t G_FLUID = 0, G_SOLID = 1, G_WALL = 2, G_CONJ = 3, G_KIND = 3, G_INLET = 16;

__device__ __forceinline__ int cx7(int i) { return i == 1 ? 1 : (i == 2 ? -1 : 0); }
__device__ __forceinline__ int cy7(int i) { return i == 3 ? 1 : (i == 4 ? -1 : 0); }
__device__ __forceinline__ int cz7(int i) { return i == 5 ? 1 : (i == 6 ? -1 : 0); }
__device__ __forceinline__ int opp7(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
__device__ __forceinline__ float w7(int i) { return i == 0 ? 0.25f : 0.125f; }

// neighbour coordinate along one axis; false when it leaves a non-periodic domain
__device__ __forceinline__ bool nb7(int c, int d, int n, int per, int& out) {
  int v = c + d;
  if (v < 0 || v >= n) {
    if (!per) return false;
    v = v < 0 ? v + n : v - n;
  }
  out = v;
  return true;
}

__global__ __launch_bounds__(kBlock) void d3q7_step_kernel(const float* __restrict__ gin, float* __restrict__ gout,
                                                           float* __restrict__ T_out, const float* __restrict__ u,
                                                           const float* __restrict__ v, const float* __restrict__ w,
                                                           const uint8_t* __restrict__ flag,
                                                           const uint8_t* __restrict__ nbmask,
                                                           const float* __restrict__ T_solid,
                                                           const float* __restrict__ Q, float* __restrict__ acc,
                                                           int nx, int ny, int nz,
                                                           int px, int py, int pz, float omega, float omega_s,
                                                           float T_in) {
  const int64_t N = (int64_t)nx * ny * nz;
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (x >= N) return;
  // N < 2^31 (checked by the entry point): 32-bit index arithmetic
  const unsigned nynz = (unsigned)ny * (unsigned)nz, ux_ = (unsigned)x;
  const int i = (int)(ux_ / nynz);
  const unsigned rem = ux_ - (unsigned)i * nynz;
  const int j = (int)(rem / (unsigned)nz), k = (int)(rem - (unsigned)j * (unsigned)nz);
  // The node's own flag and mask first: walls return and solved solids take their own
  // branch before any pull is issued (the heat sink's plate and pins are whole wavefronts
  // of solid nodes; issuing their pulls as well cost 4 %). A fluid node's seven pulls from
  // x - c_q (or, outside a non-periodic face, from x itself), the flags its mask names and
  // the velocity then issue together before the first wait.
  const uint8_t fl = flag[x];
  // bit q - 1 of nbmask: x - c_q is a wall or a solid node inside the domain; only
  // those neighbours' flags are read (without the mask, all six are). On a solved solid
  // node the bits mean instead: x + c_q is a fluid node off the inlet (see below).
  const uint32_t nbm = nbmask ? (uint32_t)nbmask[x] : 0x3Fu;
  const uint8_t kind = fl & G_KIND;
  if (kind == G_WALL) return;
  if (kind == G_SOLID) {
    // energy the fluid neighbours hand to this solid node in this step, formed from the
    // same operands and in the same order as the neighbours' anti-bounce-back inputs
    const float ts = T_solid[x];
    T_out[x] = ts;
    if (!acc) return;
    // with the mask, its bits name the fluid neighbours off the inlet, so neither their
    // flags nor (for solid nodes without fluid neighbours) anything else is read
    if (nbmask && nbm == 0u) return;
    float e = 0.0f;
#pragma unroll
    for (int q = 1; q < 7; ++q) {
      int a = i, b = j, c = k;
      if (!(nb7(i, cx7(q), nx, px, a) && nb7(j, cy7(q), ny, py, b) && nb7(k, cz7(q), nz, pz, c))) continue;
      const int64_t s = ((int64_t)a * ny + b) * nz + c;
      if (nbmask) {
        if (!((nbm >> (q - 1)) & 1u)) continue;
      } else {
        const uint8_t fs = flag[s];
        if ((fs & G_KIND) != G_FLUID || (fs & G_INLET)) continue;
      }
      const float out = gin[opp7(q) * N + s];
      const float two_wt = 2.0f * w7(q) * ts;
      const float back = two_wt - out;
      e = e + (out - back);
    }
    acc[x] = acc[x] + e;
    return;
  }
  int64_t src_node[7];
  uint8_t nk[7];
  float g[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    int a = i, b = j, c = k;
    const bool in = nb7(i, -cx7(q), nx, px, a) && nb7(j, -cy7(q), ny, py, b) && nb7(k, -cz7(q), nz, pz, c);
    src_node[q] = in ? ((int64_t)a * ny + b) * nz + c : x;
    g[q] = __builtin_nontemporal_load(gin + q * N + src_node[q]);
    nk[q] = G_FLUID;
    if (q > 0 && in && ((nbm >> (q - 1)) & 1u)) nk[q] = (uint8_t)(flag[src_node[q]] & G_KIND);
  }
  const float u0 = u[x], v0 = v[x], w0 = w[x];
  // a conducting node carried on the lattice: no velocity, its own relaxation rate
  const bool conj = kind == G_CONJ;
  const float ux = conj ? 0.0f : u0, uy = conj ? 0.0f : v0, uz = conj ? 0.0f : w0;
  const float om = conj ? omega_s : omega;
  if (fl & G_INLET) {
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const float cu = (float)cx7(q) * ux + (float)cy7(q) * uy + (float)cz7(q) * uz;
      const float gi = w7(q) * T_in * (1.0f + 4.0f * cu);
      gout[q * N + x] = gi;
    }
    T_out[x] = T_in;
    return;
  }
  // inputs from walls and solids are replaced
#pragma unroll
  for (int q = 1; q < 7; ++q) {
    if (nk[q] == G_WALL) {
      g[q] = gin[opp7(q) * N + x];
    } else if (nk[q] == G_SOLID) {
      const float two_wt = 2.0f * w7(q) * T_solid[src_node[q]];
      g[q] = two_wt - gin[opp7(q) * N + x];
    }
  }
  float T = g[0];
#pragma unroll
  for (int q = 1; q < 7; ++q) T = T + g[q];
  T_out[x] = T;
  const float src = Q ? Q[x] : 0.0f;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const float cu = (float)cx7(q) * ux + (float)cy7(q) * uy + (float)cz7(q) * uz;
    const float geq = w7(q) * T * (1.0f + 4.0f * cu);
    float o = g[q] + om * (geq - g[q]);
    if (Q) o = o + w7(q) * src;
    __builtin_nontemporal_store(o, gout + q * N + x);
  }
}

// g = g^eq(T, u) on fluid nodes (the initial state)
__global__ __launch_bounds__(kBlock) void d3q7_init_kernel(float* __restrict__ g, const float* __restrict__ T,
                                                           const float* __restrict__ u, const float* __restrict__ v,
                                                           const float* __restrict__ w, int64_t N) {
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (x >= N) return;
  const float t = T[x], ux = u[x], uy = v[x], uz = w[x];
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const float cu = (float)cx7(q) * ux + (float)cy7(q) * uy + (float)cz7(q) * uz;
    g[q * N + x] = w7(q) * t * (1.0f + 4.0f * cu);
  }
}

// Box gather/scatter of the solid conduction step. The step runs on a box that holds
// every conducting node with at least one non-conducting cell around it, origin
// (ox, oy, oz) in the flow grid (may be -1), extents (bx, by, bz); box cells outside
// the flow grid or not conducting are not unknowns (b = x0 = 0 there).
struct Box {
  int nx, ny, nz, ox, oy, oz, bx, by, bz;
};

__device__ __forceinline__ bool box_to_grid(const Box& b, int64_t t, int64_t& g) {
  const int bi = (int)(t / ((int64_t)b.by * b.bz)), bj = (int)((t / b.bz) % b.by), bk = (int)(t % b.bz);
  const int i = bi + b.ox, j = bj + b.oy, k = bk + b.oz;
  if (i < 0 || j < 0 || k < 0 || i >= b.nx || j >= b.ny || k >= b.nz) return false;
  g = ((int64_t)i * b.ny + j) * b.nz + k;
  return true;
}

// b = T_s + dt src + acc / rho_cp on conducting cells: the right-hand side of
// (I + dt L) T = b, L the solid conduction operator divided by rho_cp. Start x0 = T_s;
// with the previous interval's start temperature T_prev the linear extrapolation
// 2 T_s - T_prev, and with the one before, T_prev2, the quadratic one
// 3 (T_s - T_prev) + T_prev2 (the solid heats smoothly, so the solve starts closer to
// its answer)
__global__ __launch_bounds__(kBlock) void solid_rhs_kernel(const float* __restrict__ T_solid,
                                                           const float* __restrict__ T_prev,
                                                           const float* __restrict__ T_prev2,
                                                           const float* __restrict__ acc,
                                                           const float* __restrict__ src,
                                                           const uint8_t* __restrict__ flag, Box box, double inv_rcp,
                                                           double dt, double* __restrict__ b, double* __restrict__ x0) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)box.bx * box.by * box.bz) return;
  int64_t g;
  double bv = 0.0, xv = 0.0;
  if (box_to_grid(box, t, g) && (flag[g] & G_KIND) == G_SOLID) {
    const double ts = (double)T_solid[g];
    const double s = src ? dt * (double)src[g] : 0.0;
    bv = (ts + s) + (double)acc[g] * inv_rcp;
    if (T_prev && T_prev2)
      xv = 3.0 * (ts - (double)T_prev[g]) + (double)T_prev2[g];
    else
      xv = T_prev ? 2.0 * ts - (double)T_prev[g] : ts;
  }
  b[t] = bv;
  x0[t] = xv;
}

// T_s = x on conducting cells, and their energy accumulators restart at zero
__global__ __launch_bounds__(kBlock) void solid_store_kernel(const double* __restrict__ x, float* __restrict__ T_solid,
                                                             float* __restrict__ acc, const uint8_t* __restrict__ flag,
                                                             Box box) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)box.bx * box.by * box.bz) return;
  int64_t g;
  if (box_to_grid(box, t, g) && (flag[g] & G_KIND) == G_SOLID) {
    T_solid[g] = (float)x[t];
    acc[g] = 0.0f;
  }
}

int make_box(Box& b, int nx, int ny, int nz, int ox, int oy, int oz, int bx, int by, int bz) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0 && bx > 0 && by > 0 && bz > 0, "cht_solid: extents must be positive");
  FC_REQUIRE(ox >= -1 && oy >= -1 && oz >= -1 && ox < nx && oy < ny && oz < nz, "cht_solid: box origin out of range");
  b = Box{nx, ny, nz, ox, oy, oz, bx, by, bz};
  return 0;
}

}  // namespace

extern "C" {

int fc_d3q7_step(const float* g_in, float* g_out, float* T_out, const float* u, const float* v, const float* w,
                 const uint8_t* flag, const uint8_t* nbmask, const float* T_solid, const float* Q, float* acc, int nx,
                 int ny, int nz, int px, int py, int pz, float omega, float omega_s, float T_in, hipStream_t stream) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "d3q7_step: grid extents must be positive");
  FC_REQUIRE(g_in && g_out && T_out && u && v && w && flag && T_solid, "d3q7_step: null pointer argument");
  FC_REQUIRE(g_in != g_out, "d3q7_step: input and output populations must be different buffers");
  FC_REQUIRE(omega > 0.f && omega < 2.f && omega_s > 0.f && omega_s < 2.f,
             "d3q7_step: omega and omega_s must be in (0, 2)");
  const int64_t N = (int64_t)nx * ny * nz;
  FC_REQUIRE(N < ((int64_t)1 << 31), "d3q7_step: the grid must hold fewer than 2^31 nodes");
  hipLaunchKernelGGL(d3q7_step_kernel, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, stream, g_in, g_out, T_out, u, v,
                     w, flag, nbmask, T_solid, Q, acc, nx, ny, nz, px != 0, py != 0, pz != 0, omega, omega_s, T_in);
  return check_launch("d3q7_step");
}

int fc_d3q7_init(float* g, const float* T, const float* u, const float* v, const float* w, int64_t n,
                 hipStream_t stream) {
  FC_REQUIRE(n >= 0 && g && T && u && v && w, "d3q7_init: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(d3q7_init_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, g, T, u, v, w, n);
  return check_launch("d3q7_init");
}

int fc_cht_solid_rhs(const float* T_solid, const float* T_prev, const float* T_prev2, const float* acc,
                     const float* src, const uint8_t* flag, int nx, int ny, int nz, int ox, int oy, int oz, int bx,
                     int by, int bz, double rho_cp, double dt, double* b, double* x0, hipStream_t stream) {
  Box box;
  if (int rc = make_box(box, nx, ny, nz, ox, oy, oz, bx, by, bz)) return rc;
  FC_REQUIRE(T_solid && acc && flag && b && x0, "cht_solid_rhs: null pointer argument");
  FC_REQUIRE(rho_cp > 0.0 && dt > 0.0, "cht_solid_rhs: rho_cp and dt must be positive");
  FC_REQUIRE(!T_prev2 || T_prev, "cht_solid_rhs: T_prev2 needs T_prev");
  const int64_t n = (int64_t)bx * by * bz;
  hipLaunchKernelGGL(solid_rhs_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, T_solid, T_prev, T_prev2,
                     acc, src, flag, box, 1.0 / rho_cp, dt, b, x0);
  return check_launch("cht_solid_rhs");
}

int fc_cht_solid_store(const double* x, float* T_solid, float* acc, const uint8_t* flag, int nx, int ny, int nz,
                       int ox, int oy, int oz, int bx, int by, int bz, hipStream_t stream) {
  Box box;
  if (int rc = make_box(box, nx, ny, nz, ox, oy, oz, bx, by, bz)) return rc;
  FC_REQUIRE(x && T_solid && acc && flag, "cht_solid_store: null pointer argument");
  const int64_t n = (int64_t)bx * by * bz;
  hipLaunchKernelGGL(solid_store_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, x, T_solid, acc, flag,
                     box);
  return check_launch("cht_solid_store");
}

}  // extern "C"
