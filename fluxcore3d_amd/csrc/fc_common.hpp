// fc_common.hpp — shared definitions for the gfx950 kernels of fluxcore3d_amd.
//
// D3Q19 lattice (velocity ordering of LBM_CHT.py:535-558 / lbm_kernels.cu:59-62),
// the reference's FP32 rest weights (lbm_kernels.cu:65-71: the literals
// 0.05555556f / 0.02777778f sit 1 ulp above fp32(1/18), fp32(1/36); solid
// resets must reproduce them bit-exactly), the error channel of the C ABI and
// small launch helpers.
//
// Layout convention used by every LBM entry point: populations are SoA,
// f[q * N + flat] with flat = (i * ny + j) * nz + k (C order of (19,nx,ny,nz)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace fc {

constexpr int Q = 19;

__host__ __device__ constexpr int CX(int q) {
  constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0};
  return t[q];
}
__host__ __device__ constexpr int CY(int q) {
  constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
  return t[q];
}
__host__ __device__ constexpr int CZ(int q) {
  constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
  return t[q];
}
__host__ __device__ constexpr int OPP(int q) {
  constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  return t[q];
}
// Rest weights with the reference's literal values (see header comment).
__host__ __device__ constexpr float WQ(int q) {
  return q == 0 ? 0.33333333f : (q < 7 ? 0.05555556f : 0.02777778f);
}
__host__ __device__ constexpr int C_AX(int q, int axis) {
  return axis == 0 ? CX(q) : (axis == 1 ? CY(q) : CZ(q));
}

// D3Q19 moment basis (d'Humieres 2002 ordering, identical to the table in
// flow_solver_upgrade.py:98-120), generated from its defining polynomials in
// the lattice velocity c (|c|^2 = c2):
//   1, 19c2-30, (21c2^2-53c2+24)/2, cx, (5c2-9)cx, cy, (5c2-9)cy, cz,
//   (5c2-9)cz, 3cx^2-c2, (3c2-5)(3cx^2-c2), cy^2-cz^2, (3c2-5)(cy^2-cz^2),
//   cxcy, cycz, cxcz, cx(cy^2-cz^2), cy(cz^2-cx^2), cz(cx^2-cy^2).
// The rows are mutually orthogonal, so M^-1 = M^T diag(1/|M_k|^2).
__host__ __device__ constexpr int MROW(int k, int q) {
  const int x = CX(q), y = CY(q), z = CZ(q);
  const int c2 = x * x + y * y + z * z;
  switch (k) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return x;
    case 4: return (5 * c2 - 9) * x;
    case 5: return y;
    case 6: return (5 * c2 - 9) * y;
    case 7: return z;
    case 8: return (5 * c2 - 9) * z;
    case 9: return 3 * x * x - c2;
    case 10: return (3 * c2 - 5) * (3 * x * x - c2);
    case 11: return y * y - z * z;
    case 12: return (3 * c2 - 5) * (y * y - z * z);
    case 13: return x * y;
    case 14: return y * z;
    case 15: return x * z;
    case 16: return x * (y * y - z * z);
    case 17: return y * (z * z - x * x);
    default: return z * (x * x - y * y);
  }
}
__host__ __device__ constexpr int MNORM2(int k) {
  int s = 0;
  for (int q = 0; q < 19; ++q) s += MROW(k, q) * MROW(k, q);
  return s;
}

// Equilibrium w_i rho (1 + cu + cu^2/2 - 1.5 u^2) with cu = 3 c.u, evaluated
// in the operand order of compute_feq (lbm_kernels.cu:77-113).
__device__ __forceinline__ void feq19(float rho, float ux, float uy, float uz, float feq[19]) {
  const float t = 1.5f * (ux * ux + uy * uy + uz * uz);
  const float ox = 3.f * ux, oy = 3.f * uy, oz = 3.f * uz;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    float cu = 0.f;
    if (CX(q) != 0) cu = CX(q) > 0 ? ox : -ox;
    if (CY(q) != 0) cu = cu + (CY(q) > 0 ? oy : -oy);
    if (CZ(q) != 0) cu = cu + (CZ(q) > 0 ? oz : -oz);
    feq[q] = (WQ(q) * rho) * (1.f + cu + 0.5f * cu * cu - t);
  }
}

// rho, j = sum_q f_q (1, c_q), summed in q order (lbm_kernels.cu:286-292).
__device__ __forceinline__ void moments0(const float f[19], float& rho, float& jx, float& jy, float& jz) {
  rho = 0.f; jx = 0.f; jy = 0.f; jz = 0.f;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += f[q];
    if (CX(q) != 0) jx += CX(q) > 0 ? f[q] : -f[q];
    if (CY(q) != 0) jy += CY(q) > 0 ? f[q] : -f[q];
    if (CZ(q) != 0) jz += CZ(q) > 0 ? f[q] : -f[q];
  }
}

// MRT relaxation of the non-equilibrium part: returns df = M^-1 S M (f - feq)
// for the 15 non-conserved moments (rows 0,3,5,7 are never relaxed,
// lbm_kernels.cu:425-427). S is indexed by moment row.
__device__ __forceinline__ void mrt_df(const float fneq[19], const float S[19], float df[19]) {
  float dm[19];
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    if (k == 0 || k == 3 || k == 5 || k == 7) { dm[k] = 0.f; continue; }
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const int c = MROW(k, q);
      if (c == 1) m += fneq[q];
      else if (c == -1) m -= fneq[q];
      else if (c != 0) m += (float)c * fneq[q];
    }
    dm[k] = S[k] * m * (1.0f / (float)MNORM2(k));
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 19; ++k) {
      if (k == 0 || k == 3 || k == 5 || k == 7) continue;
      const int c = MROW(k, q);
      if (c == 1) s += dm[k];
      else if (c == -1) s -= dm[k];
      else if (c != 0) s += (float)c * dm[k];
    }
    df[q] = s;
  }
}

// Face form of the 7-point conduction coefficients. One array per axis, with the
// cell layout: slot c holds the coefficient of the face between cell c and its
// +1 neighbour along the axis. The last plane has no such face; its slot holds
// the low-boundary coefficient of the first plane where the low ghost is the cell
// itself (flow axis, inlet at 0), and is unused otherwise. Mirrored ghosts
// (high face -> n-2, low face -> 1) reuse the face towards that neighbour.
// pos = index along the axis, stride = its element stride.
__device__ __forceinline__ int face_hi(int idx, int pos, int n, int stride) {
  return pos < n - 1 ? idx : idx - stride;
}
__device__ __forceinline__ int face_lo(int idx, int pos, int n, int stride, int ghost_mirrors) {
  return pos > 0 ? idx - stride : (ghost_mirrors ? idx : idx + (n - 1) * stride);
}

// FP64 sum over a 64-wide wavefront with DPP lane moves (quad permutes, half-row
// and row mirrors), the four 16-lane rows combined through readlane. Every lane
// of the wavefront must be active; all lanes return the sum.
__device__ __forceinline__ double dpp_move(double v, int ctrl_sel) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  switch (ctrl_sel) {
    case 0:  // quad_perm [1,0,3,2]
      lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
      break;
    case 1:  // quad_perm [2,3,0,1]
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
      break;
    case 2:  // row_half_mirror
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x141, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x141, 0xF, 0xF, false);
      break;
    default:  // row_mirror
      lo = __builtin_amdgcn_update_dpp(0, lo, 0x140, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(0, hi, 0x140, 0xF, 0xF, false);
      break;
  }
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_move(v, 0);
  v += dpp_move(v, 1);
  v += dpp_move(v, 2);
  v += dpp_move(v, 3);
  // every lane of each 16-lane row now holds that row's sum
  const int lo0 = __builtin_amdgcn_readlane(__double2loint(v), 0), hi0 = __builtin_amdgcn_readlane(__double2hiint(v), 0);
  const int lo1 = __builtin_amdgcn_readlane(__double2loint(v), 16), hi1 = __builtin_amdgcn_readlane(__double2hiint(v), 16);
  const int lo2 = __builtin_amdgcn_readlane(__double2loint(v), 32), hi2 = __builtin_amdgcn_readlane(__double2hiint(v), 32);
  const int lo3 = __builtin_amdgcn_readlane(__double2loint(v), 48), hi3 = __builtin_amdgcn_readlane(__double2hiint(v), 48);
  return (__hiloint2double(hi0, lo0) + __hiloint2double(hi1, lo1)) +
         (__hiloint2double(hi2, lo2) + __hiloint2double(hi3, lo3));
}

// Round-up grid helper.
inline unsigned grid_for(int64_t n, int block) {
  return (unsigned)((n + block - 1) / block);
}

}  // namespace fc

// ---------------------------------------------------------------------------
// Error channel of the C ABI: every entry point returns 0 on success or a
// non-zero code and leaves a message readable through fc_last_error().
// ---------------------------------------------------------------------------
extern "C" const char* fc_last_error(void);
namespace fc {
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);
}

#define FC_REQUIRE(cond, ...)                                   \
  do {                                                          \
    if (!(cond)) return ::fc::set_error(1, __VA_ARGS__);        \
  } while (0)
