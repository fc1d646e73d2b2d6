// lbm_aa.hip — fused D3Q19 collide + stream + Bouzidi step, AA pattern, in place.
//
// One population buffer pop[19][Ns] (SoA over a padded grid), one thread per
// lattice node, one launch per time step:
//   even step: node x reads slot (i, x) for every i, collides, writes f*_i to
//              slot (opp i, x);
//   odd step:  node x reads slot (opp i, x - c_i), collides, writes f*_i to
//              slot (i, x + c_i).
// Every slot is read and written by exactly one node in a step, so the update
// is race-free in place and moves 19 loads + 19 stores per node (152 B FP32,
// 76 B for 16-bit storage).
//
// The step reproduces the reference step of lbm_cuda_ext.py:445-517
// (collide -> [sponge] -> stream_pull -> ibb -> reset_solid -> flow BCs):
//  * an input population whose upstream node is solid is the rest weight w_i
//    (the reference streams f_post = W out of solid nodes, lbm_kernels.cu:277);
//    it is substituted in registers, solid slots are never read;
//  * a non-periodic face keeps the node's own post-collision value
//    (lbm_kernels.cu:472-491) through a one-cell ghost layer;
//  * Bouzidi links are handled per node. For a link (x, t) with weights
//    (w1, w2) the overridden input at the next step is
//       w1 f*_{opp t}(x) + w2 f*_t(x)            (no second node), or
//       w1 f*_{opp t}(x) + w2 f*_{opp t}(x_ff)   (second fluid node x_ff).
//    The own-node part is computed by x at the end of its collision and kept
//    in A[l] (read back by the same thread next step). The second-node part
//    is published by x_ff into B[parity][m] and read by x next step (double
//    buffered by step parity). t = q for the reference's link direction
//    choice (lbm_kernels.cu:567-587); the host passes t = opp(q) for textbook
//    Bouzidi. Either way the kernel is the same.
//  * planes with inlet/outlet conditions are rebuilt by plane kernels between
//    steps (the nodes then take their inputs from plain slots).
#include "fc_common.hpp"
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

using namespace fc;

namespace {

constexpr int kBlock = 256;

// node flag bits
constexpr uint8_t F_SOLID = 1;    // flow-solid node: no populations, macros (1,0,0,0)
constexpr uint8_t F_SPECIAL = 2;  // has a record in the special-node table
constexpr int F_BC_SHIFT = 3;     // bits 3..5: 1 + index of the first BC plane holding the node

struct Grid {
  int nx, ny, nz;
  int gx, gy, gz;  // ghost width per axis (0 periodic, 1 non-periodic)
  int PY, PZ;      // padded extents along y and z
  int64_t Ns;      // padded node count = stride between population arrays
  int64_t N;       // interior node count
};

struct Tables {
  const uint8_t* nflag;      // [N]
  const int32_t* wave_base;  // [ceil(N/64)] special nodes before each 64-node group
  const int4* wave_meta;     // [ceil(N/64) + 1] {first link, first second-node link, first publish, first source}
  // links of a group: the nV second-node links first (entry e <-> B slot v_lo + e), then the others
  const int32_t* lk_code;    // [n_links] node in group | t << 6 | opp t << 11 | V << 16 | override << 17 | v << 18
  const int32_t* s_code;     // [n_src] node in group | q << 6 | override << 17
  const int32_t* f_code;     // [n_ff] node in group | p << 6
  const int32_t* lane_nodes;  // [n_lanes] node of each lane (null: lane r is node r)
  int64_t n_lanes;
  const int4* sp_mask;       // [n_special] {L, S, V, F}
  const int4* sp_base;       // [n_special] {first plain link, v_base, ff_base, first second-node link}
  const float* w1;           // [n_links]
  const float* w2;           // [n_links], or null: w2 = 1 - w1 (the reference's weights, exactly)
  float* A;                  // [n_links]
  float* B;                  // [2][n_v]
  const int32_t* ff_idx;     // [n_ff], or null: f_code bits 11.. hold the slot relative to the group's v_lo
  int64_t n_v;
  // compact storage (null: dense padded storage): per row R = I * PY + J and word
  // w = K / 32 of the padded grid, {presence mask, present nodes before the word}
  const uint2* row_meta;
  int wpr;                   // words per row
};

struct Collision {
  int kind;          // 0 BGK, 1 TRT, 2 MRT, 3 MRT + Smagorinsky
  float om1, om2;    // BGK omega | TRT omega+, omega- | Smagorinsky base omega0
  float S[19];       // MRT rates by moment row
  float cs_dx2;      // (Cs dx)^2 for Smagorinsky
  float F[3];        // body force (lattice units), zero when unused
  int use_force;
  const float* sponge;  // sigma profile along sponge_axis, or null
  int sponge_axis;
};

// ---- storage types: FP32 values, or 16-bit deviations f - w_i (rho0 = 1) --
// ld/st: plain accesses; ld_nt/st_nt: non-temporal (streaming) accesses for the
// step's population sweep, which touches every slot once per step.
template <typename ST> struct Store;
template <> struct Store<float> {
  static __device__ __forceinline__ float ld(const float* p, int) { return *p; }
  static __device__ __forceinline__ void st(float* p, float v, int) { *p = v; }
  static __device__ __forceinline__ float ld_nt(const float* p, int) { return __builtin_nontemporal_load(p); }
  static __device__ __forceinline__ void st_nt(float* p, float v, int) { __builtin_nontemporal_store(v, p); }
  // the value a store of v followed by a load returns
  static __device__ __forceinline__ float rt(float v, int) { return v; }
};
template <> struct Store<__half> {
  static __device__ __forceinline__ float ld(const __half* p, int q) { return __half2float(*p) + WQ(q); }
  static __device__ __forceinline__ void st(__half* p, float v, int q) { *p = __float2half(v - WQ(q)); }
  static __device__ __forceinline__ float ld_nt(const __half* p, int q) {
    const unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
    return __half2float(__ushort_as_half(b)) + WQ(q);
  }
  static __device__ __forceinline__ void st_nt(__half* p, float v, int q) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(v - WQ(q))), reinterpret_cast<unsigned short*>(p));
  }
  static __device__ __forceinline__ float rt(float v, int q) { return __half2float(__float2half(v - WQ(q))) + WQ(q); }
};
template <> struct Store<__hip_bfloat16> {
  static __device__ __forceinline__ float ld(const __hip_bfloat16* p, int q) {
    return __bfloat162float(*p) + WQ(q);
  }
  static __device__ __forceinline__ void st(__hip_bfloat16* p, float v, int q) {
    *p = __float2bfloat16(v - WQ(q));
  }
  static __device__ __forceinline__ float ld_nt(const __hip_bfloat16* p, int q) {
    const unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
    return __bfloat162float(__ushort_as_bfloat16(b)) + WQ(q);
  }
  static __device__ __forceinline__ void st_nt(__hip_bfloat16* p, float v, int q) {
    __builtin_nontemporal_store(__bfloat16_as_ushort(__float2bfloat16(v - WQ(q))),
                                reinterpret_cast<unsigned short*>(p));
  }
  static __device__ __forceinline__ float rt(float v, int q) {
    return __bfloat162float(__float2bfloat16(v - WQ(q))) + WQ(q);
  }
};

__device__ __forceinline__ int wrap_or_pad(int c, int d, int n, int g) {
  int v = c + d;
  if (g) return v + 1;  // ghost layer: -1 -> 0, n -> n+1
  return v < 0 ? v + n : (v >= n ? v - n : v);
}
// Storage index of the node at offset (dx, dy, dz) from interior node (i, j, k):
// its padded flat index (dense storage), or its rank among the nodes that own
// slots (compact storage). Returns false for a node without slots (compact
// storage only): it streams w_q and receives nothing.
__device__ __forceinline__ bool node_storage(const Grid& g, const Tables& t, int i, int j, int k, int dx, int dy,
                                             int dz, int64_t& s) {
  const int I = wrap_or_pad(i, dx, g.nx, g.gx);
  const int J = wrap_or_pad(j, dy, g.ny, g.gy);
  const int K = wrap_or_pad(k, dz, g.nz, g.gz);
  if (!t.row_meta) {
    s = ((int64_t)I * g.PY + J) * g.PZ + K;
    return true;
  }
  const uint2 m = t.row_meta[((int64_t)I * g.PY + J) * t.wpr + (K >> 5)];
  const uint32_t b = (uint32_t)K & 31u;
  s = (int64_t)m.y + __popc(m.x & ((1u << b) - 1u));
  return ((m.x >> b) & 1u) != 0u;
}
// true when x - c_q leaves the domain through a non-periodic face
__device__ __forceinline__ bool oob_source(const Grid& g, int i, int j, int k, int q) {
  return (g.gx && (i - CX(q) < 0 || i - CX(q) >= g.nx)) || (g.gy && (j - CY(q) < 0 || j - CY(q) >= g.ny)) ||
         (g.gz && (k - CZ(q) < 0 || k - CZ(q) >= g.nz));
}

template <int BIT>
__device__ __forceinline__ int rank_below(uint32_t mask) {
  return __popc(mask & ((1u << BIT) - 1u));
}

// Plain slot that holds input population q of node (i,j,k) at this parity. With
// compact storage a slot of a node without slots reads as w_q and is not written.
template <typename ST>
__device__ __forceinline__ float load_plain(const ST* pop, const Grid& g, const Tables& t, int parity, int i, int j,
                                            int k, int q) {
  int64_t s;
  const int d = parity == 0 ? 0 : -1;
  if (!node_storage(g, t, i, j, k, d * CX(q), d * CY(q), d * CZ(q), s)) return WQ(q);
  const int dir = parity == 0 ? q : OPP(q);
  return Store<ST>::ld(pop + (int64_t)dir * g.Ns + s, dir);
}
template <typename ST>
__device__ __forceinline__ void store_plain(ST* pop, const Grid& g, const Tables& t, int parity, int i, int j, int k,
                                            int q, float v) {
  int64_t s;
  const int d = parity == 0 ? 0 : -1;
  if (!node_storage(g, t, i, j, k, d * CX(q), d * CY(q), d * CZ(q), s)) return;
  const int dir = parity == 0 ? q : OPP(q);
  Store<ST>::st(pop + (int64_t)dir * g.Ns + s, v, dir);
}

// Second Bouzidi weight of link l: stored, or 1 - w1 (how the reference forms it in
// both branches, sdf_tracing.py:188-205, and in the fallback w1 = 1, w2 = 0).
__device__ __forceinline__ float w2_of(const Tables& t, int64_t l, float w1) { return t.w2 ? t.w2[l] : 1.f - w1; }

// All 19 inputs of a node: plain slots, rest weights for solid sources and
// Bouzidi values for links. `first` marks the step right after loading: A then
// holds the loaded populations themselves and solid sources have not streamed
// yet, so those inputs are the loaded values.
template <typename ST>
__device__ __forceinline__ void compose_inputs(const ST* __restrict__ pop, const Grid& g, const Tables& t,
                                               int parity, int first, int i, int j, int k, bool special,
                                               bool plain, int4 m, int4 b, float f[19], uint32_t want = 0x7FFFFu) {
  const uint32_t L = plain ? 0u : (uint32_t)m.x, S = plain ? 0u : (uint32_t)m.y, V = (uint32_t)m.z;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    if (!((want >> q) & 1u)) {  // an input the caller does not use
      f[q] = 0.f;
    } else if (special && ((L >> q) & 1u)) {
      const bool isv = (V >> q) & 1u;
      const int64_t l = isv ? b.w + __popc(V & ((1u << q) - 1u)) : b.x + __popc(L & ~V & ((1u << q) - 1u));
      float v = t.A[l];
      if (!first && isv) {
        const int64_t mv = b.y + __popc(V & ((1u << q) - 1u));
        v = v + w2_of(t, l, t.w1[l]) * t.B[(int64_t)(parity ^ 1) * t.n_v + mv];
      }
      f[q] = v;
    } else if (special && !first && ((S >> q) & 1u)) {
      f[q] = WQ(q);  // after the first step a solid source delivers the rest weight
    } else {
      f[q] = load_plain(pop, g, t, parity, i, j, k, q);
    }
  }
}

// Collision in registers; returns the macros of the incoming state.
// Optional Guo body force (u = (j + F/2)/rho, source (1 - s/2) F_i with the
// relaxation rate of each mode) and sponge relaxation towards feq.
__device__ __forceinline__ void guo_source(const Collision& c, float ux, float uy, float uz, float src[19]) {
  const float uf = ux * c.F[0] + uy * c.F[1] + uz * c.F[2];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const float cx = (float)CX(q), cy = (float)CY(q), cz = (float)CZ(q);
    const float cu = cx * ux + cy * uy + cz * uz;
    const float cf = cx * c.F[0] + cy * c.F[1] + cz * c.F[2];
    src[q] = WQ(q) * (3.f * (cf - uf) + 9.f * cu * cf);
  }
}

template <int KIND>
__device__ __forceinline__ void collide(const Collision& c, float f[19], float& rho, float& ux, float& uy, float& uz,
                                        float sigma) {
  float jx, jy, jz;
  moments0(f, rho, jx, jy, jz);
  if (c.use_force) {
    jx += 0.5f * c.F[0]; jy += 0.5f * c.F[1]; jz += 0.5f * c.F[2];
  }
  const float inv = 1.f / (rho + 1e-12f);
  ux = jx * inv; uy = jy * inv; uz = jz * inv;
  float feq[19];
  feq19(rho, ux, uy, uz, feq);
  if constexpr (KIND == 0) {
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = f[q] + c.om1 * (feq[q] - f[q]);
    if (c.use_force) {
      float src[19];
      guo_source(c, ux, uy, uz, src);
#pragma unroll
      for (int q = 0; q < 19; ++q) f[q] += (1.f - 0.5f * c.om1) * src[q];
    }
  } else if constexpr (KIND == 1) {
    float o[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const int p = OPP(q);
      const float fp = 0.5f * (f[q] + f[p]), fm = 0.5f * (f[q] - f[p]);
      const float ep = 0.5f * (feq[q] + feq[p]), em = 0.5f * (feq[q] - feq[p]);
      o[q] = (fp + c.om1 * (ep - fp)) + (fm + c.om2 * (em - fm));
    }
    if (c.use_force) {
      float src[19];
      guo_source(c, ux, uy, uz, src);
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const int p = OPP(q);
        o[q] += (1.f - 0.5f * c.om1) * (0.5f * (src[q] + src[p])) + (1.f - 0.5f * c.om2) * (0.5f * (src[q] - src[p]));
      }
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = o[q];
  } else {
    float S[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) S[k] = c.S[k];
    float fneq[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) fneq[q] = f[q] - feq[q];
    if constexpr (KIND == 3) {
      float pxx = 0.f, pyy = 0.f, pzz = 0.f, pxy = 0.f, pxz = 0.f, pyz = 0.f;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const float x = (float)CX(q), y = (float)CY(q), z = (float)CZ(q);
        pxx += fneq[q] * x * x; pyy += fneq[q] * y * y; pzz += fneq[q] * z * z;
        pxy += fneq[q] * x * y; pxz += fneq[q] * x * z; pyz += fneq[q] * y * z;
      }
      const float pif = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.f * (pxy * pxy + pxz * pxz + pyz * pyz));
      const float nu0 = (1.f / c.om1 - 0.5f) / 3.f;
      const float disc = nu0 * nu0 + 4.f * c.cs_dx2 * (pif * inv) * 4.5f;
      const float nu_eff = 0.5f * (nu0 + sqrtf(fmaxf(disc, nu0 * nu0)));
      const float om = fminf(1.99f, 1.f / (3.f * nu_eff + 0.5f));
      S[9] = om; S[11] = om; S[13] = om; S[14] = om; S[15] = om;
    }
    float df[19];
    mrt_df(fneq, S, df);
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = f[q] - df[q];
    if (c.use_force) {
      // f += M^-1 (I - S/2) M src: non-conserved rows through mrt_df with
      // rates -(1 - S_k/2); the momentum rows (S = 0, |M_k|^2 = 10) carry F.
      float src[19], Sm[19];
      guo_source(c, ux, uy, uz, src);
#pragma unroll
      for (int k = 0; k < 19; ++k) Sm[k] = 0.5f * S[k] - 1.f;
      mrt_df(src, Sm, df);
#pragma unroll
      for (int q = 0; q < 19; ++q)
        f[q] = f[q] - df[q] + 0.1f * ((float)CX(q) * c.F[0] + (float)CY(q) * c.F[1] + (float)CZ(q) * c.F[2]);
    }
  }
  if (sigma > 0.f) {  // sponge (lbm_kernels.cu:505-528) on the post-collision state
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = f[q] - sigma * (f[q] - feq[q]);
  }
}

__device__ __forceinline__ unsigned logical_block(unsigned bid, unsigned nblk) {
  // Consecutive node blocks on the same XCD (blocks are dealt round-robin to
  // the 8 XCDs), so stencil neighbours share that XCD's L2.
  const unsigned xcd = bid & 7u, local = bid >> 3;
  const unsigned base = nblk >> 3, rem = nblk & 7u;
  return xcd * base + min(xcd, rem) + local;
}

// MRT: non-equilibrium moments dm = S (M f - m_eq) with m_eq in closed form
// for the second-order equilibrium (the moments of compute_feq's polynomial),
// Guo source moments (1 - s/2) M F_i folded into dm, and f* = f - M^T
// diag(1/|M_k|^2) dm. Subtracting a small correction from f (rather than
// rebuilding f from its moments) keeps FP32 round-off at the level of the
// reference's f - M^-1 dm (lbm_kernels.cu:415-435).
__device__ __forceinline__ void mrt_moment_collide(const Collision& c, const float S[19], float f[19], float rho,
                                                   float ux, float uy, float uz, bool force) {
  const float u2 = ux * ux + uy * uy + uz * uz;
  const float pxx = 3.f * ux * ux - u2, pww = uy * uy - uz * uz;
  float dm[19];
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    if (k == 0 || k == 3 || k == 5 || k == 7) { dm[k] = 0.f; continue; }
    float a = 0.f;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const int cf = MROW(k, q);
      if (cf == 1) a += f[q];
      else if (cf == -1) a -= f[q];
      else if (cf != 0) a += (float)cf * f[q];
    }
    float meq;
    switch (k) {
      case 1: meq = rho * (-11.f + 19.f * u2); break;
      case 2: meq = rho * (3.f - 5.5f * u2); break;
      case 4: meq = (-2.f / 3.f) * rho * ux; break;
      case 6: meq = (-2.f / 3.f) * rho * uy; break;
      case 8: meq = (-2.f / 3.f) * rho * uz; break;
      case 9: meq = rho * pxx; break;
      case 10: meq = -0.5f * rho * pxx; break;
      case 11: meq = rho * pww; break;
      case 12: meq = -0.5f * rho * pww; break;
      case 13: meq = rho * ux * uy; break;
      case 14: meq = rho * uy * uz; break;
      case 15: meq = rho * ux * uz; break;
      default: meq = 0.f; break;
    }
    dm[k] = S[k] * (a - meq);
  }
  if (force) {
    const float Fx = c.F[0], Fy = c.F[1], Fz = c.F[2];
    const float uf = ux * Fx + uy * Fy + uz * Fz;
    const float fxx = 2.f * (2.f * ux * Fx - uy * Fy - uz * Fz), fww = 2.f * (uy * Fy - uz * Fz);
    float Fm[19];
    Fm[0] = 0.f; Fm[1] = 38.f * uf; Fm[2] = -11.f * uf;
    Fm[3] = Fx; Fm[4] = (-2.f / 3.f) * Fx; Fm[5] = Fy; Fm[6] = (-2.f / 3.f) * Fy; Fm[7] = Fz; Fm[8] = (-2.f / 3.f) * Fz;
    Fm[9] = fxx; Fm[10] = -0.5f * fxx; Fm[11] = fww; Fm[12] = -0.5f * fww;
    Fm[13] = ux * Fy + uy * Fx; Fm[14] = uy * Fz + uz * Fy; Fm[15] = ux * Fz + uz * Fx;
    Fm[16] = 0.f; Fm[17] = 0.f; Fm[18] = 0.f;
#pragma unroll
    for (int k = 0; k < 19; ++k) dm[k] -= (1.f - 0.5f * S[k]) * Fm[k];
  }
#pragma unroll
  for (int k = 0; k < 19; ++k) dm[k] *= 1.0f / (float)MNORM2(k);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    float a = 0.f;
#pragma unroll
    for (int k = 0; k < 19; ++k) {
      if (!force && (k == 0 || k == 3 || k == 5 || k == 7)) continue;
      const int cf = MROW(k, q);
      if (cf == 1) a += dm[k];
      else if (cf == -1) a -= dm[k];
      else if (cf != 0) a += (float)cf * dm[k];
    }
    f[q] -= a;
  }
}

template <int KIND>
__device__ __forceinline__ void collide_fast(const Collision& c, float f[19], float& rho, float& ux, float& uy,
                                             float& uz, float sigma) {
  if constexpr (KIND == 2) {
    float jx, jy, jz;
    moments0(f, rho, jx, jy, jz);
    const bool force = c.use_force != 0;
    if (force) { jx += 0.5f * c.F[0]; jy += 0.5f * c.F[1]; jz += 0.5f * c.F[2]; }
    const float inv = 1.f / (rho + 1e-12f);
    ux = jx * inv; uy = jy * inv; uz = jz * inv;
    float feq_s[19];
    if (sigma > 0.f) feq19(rho, ux, uy, uz, feq_s);
    mrt_moment_collide(c, c.S, f, rho, ux, uy, uz, force);
    if (sigma > 0.f) {
#pragma unroll
      for (int q = 0; q < 19; ++q) f[q] = f[q] - sigma * (f[q] - feq_s[q]);
    }
  } else {
    collide<KIND>(c, f, rho, ux, uy, uz, sigma);
  }
}

// Bouzidi inputs and outputs of a wavefront go through LDS: the 64 lanes
// deposit their 19 inputs (then their 19 post-collision values) in a 19 x 64
// tile, and the wave walks its node-major entry lists with coalesced loads,
// one entry per lane: link values A (+ w2 B of the second node) and solid-source
// constants overwrite tile inputs; new A values and second-node publications
// are read from the post-collision tile. No per-lane loop over directions.
constexpr int kWaves = kBlock / 64;
// Plane kernels (inlet/outlet conditions) run one wavefront per block: a plane holds
// tens of thousands of nodes, and single-wavefront blocks spread them over every CU.
constexpr int kPlaneBlock = 64;
constexpr int kCodeV = 1 << 16, kCodeOvr = 1 << 17;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename ST, int KIND, bool CMP>
__global__ __launch_bounds__(kBlock, (KIND == 3 ? 1 : 4)) void aa_step_kernel(ST* __restrict__ pop, Grid g, Tables t, Collision c,
                                                         int parity, int first, float* __restrict__ rho_o,
                                                         float* __restrict__ u_o, float* __restrict__ v_o,
                                                         float* __restrict__ w_o) {
  __shared__ float tile[kWaves][19][64];
  // lane r works on node x (every node, or the fluid nodes only)
  const unsigned lb = logical_block(blockIdx.x, gridDim.x);
  const int64_t r = (int64_t)lb * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool live = r < t.n_lanes;
  // per-wave entry ranges (scalar loads; waves past the last lane get empty ranges)
  const int gw = __builtin_amdgcn_readfirstlane((int)(r >> 6));
  const int last_w = (int)((t.n_lanes - 1) >> 6);
  const int gwc = min(gw, last_w);
  const int4 wm0 = t.wave_meta[gwc];
  int4 wm1 = t.wave_meta[gwc + 1];
  if (gw > last_w) wm1 = wm0;
  const int64_t x = live ? (t.lane_nodes ? (int64_t)t.lane_nodes[r] : r) : 0;

  // Populations first. Dense storage, one lane per node: solid lanes load their
  // slots too and discard them: with every lane of a wavefront active, each load
  // instruction covers whole 128-byte lines, which measured faster than the same
  // lines requested by the fluid lanes alone. Compact storage: the lanes are the
  // nodes owning slots, so a wavefront's own slots are 64 consecutive ones.
  int i = 0, j = 0, k = 0;
  uint32_t self = 0;
  float f[19];
  const uint32_t PZ = (uint32_t)g.PZ, PYZ = (uint32_t)g.PY * (uint32_t)g.PZ;
  int Im = 0, I0 = 0, Ip = 0, Jm = 0, J0 = 0, Jp = 0, Km = 0, K0 = 0, Kp = 0;
  auto slot = [&](int dx, int dy, int dz) -> uint32_t {
    const int I = dx < 0 ? Im : (dx > 0 ? Ip : I0);
    const int J = dy < 0 ? Jm : (dy > 0 ? Jp : J0);
    return __umul24((uint32_t)I, PYZ) + __umul24((uint32_t)J, PZ) + (uint32_t)(dz < 0 ? Km : (dz > 0 ? Kp : K0));
  };
  // Slot of direction q's access in this step's parity: the even step uses the
  // node's own slots, so its -1/+1 neighbour coordinates collapse onto the node's.
  // One load loop and one store loop serve both parities (no branch; w_q = w_opp(q)).
  // The x and y terms of the three coordinates are formed once per lane (24-bit
  // multiplies: padded extents stay below 2^12, PY * PZ below 2^24).
  uint32_t oI[3] = {0, 0, 0}, oJ[3] = {0, 0, 0}, oK[3] = {0, 0, 0};
  auto pslot = [&](int dx, int dy, int dz) -> uint32_t { return oI[dx + 1] + oJ[dy + 1] + oK[dz + 1]; };
  // a direction's slots: uniform base pointer plus a 32-bit byte offset per lane
  auto at = [&](int dir, uint32_t slot_idx) -> ST* {
    return (ST*)((char*)(pop + (size_t)dir * (size_t)g.Ns) + slot_idx * (uint32_t)sizeof(ST));
  };
#pragma unroll
  for (int q = 0; q < 19; ++q) f[q] = 0.f;
  // compact storage: presence (bit q) and storage index of x - c_q, the node that
  // wrote input q (q = 0: the node itself); x + c_q is x - c_opp(q)
  uint32_t pres = 0;
  uint32_t nb[19];
  const bool own_lines = (g.gx | g.gy | g.gz) == 0 && t.lane_nodes != nullptr;  // lane r owns slot r
#pragma unroll
  for (int q = 0; q < 19; ++q) nb[q] = 0;
  if (live) {
    const uint32_t xu = (uint32_t)x;
    const uint32_t nynz = (uint32_t)g.ny * (uint32_t)g.nz;
    i = (int)(xu / nynz);
    const uint32_t rem = xu - (uint32_t)i * nynz;
    j = (int)(rem / (uint32_t)g.nz);
    k = (int)(rem - (uint32_t)j * (uint32_t)g.nz);
    Im = wrap_or_pad(i, -1, g.nx, g.gx); I0 = i + g.gx; Ip = wrap_or_pad(i, 1, g.nx, g.gx);
    Jm = wrap_or_pad(j, -1, g.ny, g.gy); J0 = j + g.gy; Jp = wrap_or_pad(j, 1, g.ny, g.gy);
    Km = wrap_or_pad(k, -1, g.nz, g.gz); K0 = k + g.gz; Kp = wrap_or_pad(k, 1, g.nz, g.gz);
    oI[1] = __umul24((uint32_t)I0, PYZ); oJ[1] = __umul24((uint32_t)J0, PZ); oK[1] = (uint32_t)K0;
    self = oI[1] + oJ[1] + oK[1];
    oI[0] = parity ? __umul24((uint32_t)Im, PYZ) : oI[1]; oI[2] = parity ? __umul24((uint32_t)Ip, PYZ) : oI[1];
    oJ[0] = parity ? __umul24((uint32_t)Jm, PZ) : oJ[1]; oJ[2] = parity ? __umul24((uint32_t)Jp, PZ) : oJ[1];
    oK[0] = parity ? (uint32_t)Km : oK[1]; oK[2] = parity ? (uint32_t)Kp : oK[1];
    if constexpr (!CMP) {
#pragma unroll
      for (int q = 0; q < 19; ++q)
        f[q] = Store<ST>::ld_nt(at(q + parity * (OPP(q) - q), pslot(-CX(q), -CY(q), -CZ(q))), q);
    } else {
      const uint32_t wpr = (uint32_t)t.wpr;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const int dx = -CX(q), dy = -CY(q), dz = -CZ(q);
        const uint32_t I = (uint32_t)(dx < 0 ? Im : (dx > 0 ? Ip : I0));
        const uint32_t J = (uint32_t)(dy < 0 ? Jm : (dy > 0 ? Jp : J0));
        const uint32_t K = (uint32_t)(dz < 0 ? Km : (dz > 0 ? Kp : K0));
        const uint2 m = t.row_meta[(__umul24(I, (uint32_t)g.PY) + J) * wpr + (K >> 5)];
        const uint32_t b = K & 31u;
        pres |= ((m.x >> b) & 1u) << q;
        nb[q] = m.y + __popc(m.x & ((1u << b) - 1u));
      }
      // Nodes with slots load their inputs (an input written by a node without
      // slots is w_q after the first step); the others have nothing to load.
      // Temporal accesses where a wavefront's slots can start anywhere in a cache
      // line (odd steps; even steps with ghost slots among the lanes' slots): such
      // lines are shared with the neighbouring wavefronts and merged in L2. Without
      // a ghost layer lane r owns slot r, so on even steps a wavefront's own slots
      // are whole lines that no other wavefront touches: non-temporal.
      if ((pres & 1u) && parity == 0) {
        if (own_lines) {
#pragma unroll
          for (int q = 0; q < 19; ++q) f[q] = (((pres >> q) & 1u) || first) ? Store<ST>::ld_nt(at(q, nb[0]), q) : WQ(q);
        } else {
#pragma unroll
          for (int q = 0; q < 19; ++q) f[q] = (((pres >> q) & 1u) || first) ? Store<ST>::ld(at(q, nb[0]), q) : WQ(q);
        }
      } else if (pres & 1u) {
#pragma unroll
        for (int q = 0; q < 19; ++q) f[q] = ((pres >> q) & 1u) ? Store<ST>::ld(at(OPP(q), nb[q]), q) : WQ(q);
      }
    }
  }

  // Each lane's first entry of every list (and second link entry), requested
  // right behind the populations: no entry load waits for another load (the B
  // slot of second-node link e < nV is v_lo + e), so all latencies overlap.
  // Further entries are loaded in the loops below.
  const int l_lo = wm0.x, nL = wm1.x - wm0.x, v_lo = wm0.y, nV = wm1.y - wm0.y, f_lo = wm0.z, nF = wm1.z - wm0.z;
  const int s_lo = wm0.w, nSrc = wm1.w - wm0.w;
  const bool special_wave = (nL | nF | nSrc) != 0;
  const float* Bprev = t.B + (int64_t)(parity ^ 1) * t.n_v + v_lo;
  const bool hasL0 = lane < nL, hasS0 = !first && lane < nSrc, hasF0 = lane < nF, hasL1 = lane + 64 < nL;
  const bool hasB0 = !first && lane < nV, hasB1 = !first && lane + 64 < nV;
  int c0 = 0, sc0 = 0, fc0 = 0, c1 = 0;
  uint32_t ff0 = 0;  // unsigned: widening it to an address needs no wait for the load
  float a0 = 0.f, w10 = 0.f, w20 = 0.f, a1 = 0.f, w11 = 0.f, w21 = 0.f, b0 = 0.f, b1 = 0.f;
  if (hasL0) {
    c0 = t.lk_code[l_lo + lane];
    a0 = t.A[l_lo + lane];
    w10 = t.w1[l_lo + lane];
    if (t.w2) w20 = t.w2[l_lo + lane];  // else 1 - w10, formed where it is used
  }
  if (hasB0) b0 = Bprev[lane];
  if (hasL1) {
    c1 = t.lk_code[l_lo + lane + 64];
    a1 = t.A[l_lo + lane + 64];
    w11 = t.w1[l_lo + lane + 64];
    if (t.w2) w21 = t.w2[l_lo + lane + 64];
  }
  if (hasB1) b1 = Bprev[lane + 64];
  if (hasS0) sc0 = t.s_code[s_lo + lane];
  if (hasF0) {
    fc0 = t.f_code[f_lo + lane];
    if (t.ff_idx) ff0 = (uint32_t)t.ff_idx[f_lo + lane];  // else from fc0, formed at the store
  }
  const uint8_t nfl = t.nflag[x];  // x = 0 on dead lanes: no branch around the load
  const bool fluid = live && !(nfl & F_SOLID);

  // Bouzidi / solid-source inputs (wave-uniform branch; every lane takes part)
  if (special_wave) {
#pragma unroll
    for (int q = 0; q < 19; ++q) tile[wv][q][lane] = f[q];
    wave_lds_sync();
    if (hasL0 && (c0 & kCodeOvr))
      tile[wv][(c0 >> 6) & 31][c0 & 63] = hasB0 ? a0 + (t.w2 ? w20 : 1.f - w10) * b0 : a0;
    if (hasL1 && (c1 & kCodeOvr))
      tile[wv][(c1 >> 6) & 31][c1 & 63] = hasB1 ? a1 + (t.w2 ? w21 : 1.f - w11) * b1 : a1;
    for (int e = lane + 128; e < nL; e += 64) {
      const int code = t.lk_code[l_lo + e];
      float v = t.A[l_lo + e];
      if (!first && e < nV) v = v + w2_of(t, l_lo + e, t.w2 ? 0.f : t.w1[l_lo + e]) * Bprev[e];
      if (code & kCodeOvr) tile[wv][(code >> 6) & 31][code & 63] = v;
    }
    if (!first) {
      if (hasS0 && (sc0 & kCodeOvr)) tile[wv][(sc0 >> 6) & 31][sc0 & 63] = WQ((sc0 >> 6) & 31);
      for (int e = lane + 64; e < nSrc; e += 64) {
        const int code = t.s_code[s_lo + e];
        const int q = (code >> 6) & 31;
        if (code & kCodeOvr) tile[wv][q][code & 63] = WQ(q);
      }
    }
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = tile[wv][q][lane];
  }

  float rho = 1.f, ux = 0.f, uy = 0.f, uz = 0.f;
  if (fluid) {
    float sigma = 0.f;
    if (c.sponge) sigma = c.sponge[c.sponge_axis == 0 ? i : (c.sponge_axis == 1 ? j : k)];
    collide_fast<KIND>(c, f, rho, ux, uy, uz, sigma);
  } else {
    // A solid node streams the rest state w_q (the reference's f_post = W,
    // lbm_kernels.cu:277-283) into the slots its neighbours read next step, so
    // their inputs from solid nodes need no substitution; macros (1, 0, 0, 0).
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = WQ(q);
  }
  if (live) {
    if (rho_o) {
      __builtin_nontemporal_store(rho, rho_o + x);
      __builtin_nontemporal_store(ux, u_o + x);
      __builtin_nontemporal_store(uy, v_o + x);
      __builtin_nontemporal_store(uz, w_o + x);
    }
    if constexpr (!CMP) {
#pragma unroll
      for (int q = 0; q < 19; ++q)
        Store<ST>::st_nt(at(OPP(q) + parity * (q - OPP(q)), pslot(CX(q), CY(q), CZ(q))), f[q], q);
    } else if ((pres & 1u) && parity == 0) {
      // even: the node's own slots
      if (own_lines) {
#pragma unroll
        for (int q = 0; q < 19; ++q) Store<ST>::st_nt(at(OPP(q), nb[0]), f[q], q);
      } else {
#pragma unroll
        for (int q = 0; q < 19; ++q) Store<ST>::st(at(OPP(q), nb[0]), f[q], q);
      }
    } else if (pres & 1u) {
      // odd: slots of x + c_q, where that node has slots
#pragma unroll
      for (int q = 0; q < 19; ++q)
        if ((pres >> OPP(q)) & 1u) Store<ST>::st(at(q, nb[OPP(q)]), f[q], q);
    }
    // copy-self on non-periodic faces through the ghost layer
    if (fluid && ((g.gx && (i == 0 || i == g.nx - 1)) || (g.gy && (j == 0 || j == g.ny - 1)) ||
                  (g.gz && (k == 0 || k == g.nz - 1)))) {
#pragma unroll
      for (int q = 1; q < 19; ++q) {
        if (!oob_source(g, i, j, k, q)) continue;
        if constexpr (!CMP) {
          if (parity == 0)
            Store<ST>::st(pop + (size_t)OPP(q) * (size_t)g.Ns + slot(-CX(q), -CY(q), -CZ(q)), f[q], OPP(q));
          else
            Store<ST>::st(pop + (size_t)q * (size_t)g.Ns + self, f[q], q);
        } else {
          if (parity == 0)
            Store<ST>::st(pop + (size_t)OPP(q) * (size_t)g.Ns + nb[q], f[q], OPP(q));
          else
            Store<ST>::st(pop + (size_t)q * (size_t)g.Ns + nb[0], f[q], q);
        }
      }
    }
  }

  // own-node parts of the next link values and second-node publications
  if (special_wave) {
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < 19; ++q) tile[wv][q][lane] = f[q];
    wave_lds_sync();
    if (hasL0) {
      const int n = c0 & 63;
      float a = w10 * tile[wv][(c0 >> 11) & 31][n];
      if (!(c0 & kCodeV)) a = a + (t.w2 ? w20 : 1.f - w10) * tile[wv][(c0 >> 6) & 31][n];
      t.A[l_lo + lane] = a;
    }
    if (hasL1) {
      const int n = c1 & 63;
      float a = w11 * tile[wv][(c1 >> 11) & 31][n];
      if (!(c1 & kCodeV)) a = a + (t.w2 ? w21 : 1.f - w11) * tile[wv][(c1 >> 6) & 31][n];
      t.A[l_lo + lane + 64] = a;
    }
    for (int e = lane + 128; e < nL; e += 64) {
      const int code = t.lk_code[l_lo + e];
      const int n = code & 63;
      const float w1e = t.w1[l_lo + e];
      float a = w1e * tile[wv][(code >> 11) & 31][n];
      if (!(code & kCodeV)) a = a + w2_of(t, l_lo + e, w1e) * tile[wv][(code >> 6) & 31][n];
      t.A[l_lo + e] = a;
    }
    float* Bw = t.B + (int64_t)parity * t.n_v;
    if (hasF0) Bw[t.ff_idx ? (size_t)ff0 : (size_t)(v_lo + (fc0 >> 11))] = tile[wv][(fc0 >> 6) & 31][fc0 & 63];
    for (int e = lane + 64; e < nF; e += 64) {
      const int code = t.f_code[f_lo + e];
      Bw[t.ff_idx ? t.ff_idx[f_lo + e] : v_lo + (code >> 11)] = tile[wv][(code >> 6) & 31][code & 63];
    }
  }
}

// macros of solid nodes: rho = 1, u = 0 (Q5)
__global__ __launch_bounds__(kBlock) void aa_solid_macros_kernel(const uint8_t* __restrict__ nflag, int64_t N,
                                                                 float* __restrict__ rho, float* __restrict__ u,
                                                                 float* __restrict__ v, float* __restrict__ w) {
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (x >= N || !(nflag[x] & F_SOLID)) return;
  rho[x] = 1.f; u[x] = 0.f; v[x] = 0.f; w[x] = 0.f;
}

// ---- helpers for kernels that visit arbitrary nodes -------------------------
__device__ __forceinline__ int64_t special_index_scan(const Tables& t, int64_t x) {
  // special nodes before x in its 64-node group: the group's whole 8-byte words of
  // flags below x (every byte of them lies below x, inside the table) by a popcount
  // of their special bits, the rest byte by byte
  int64_t s = t.wave_base[x >> 6];
  const int64_t base = x & ~(int64_t)63;
  const int n = (int)(x - base);
  int p = 0;
  if ((reinterpret_cast<uintptr_t>(t.nflag) & 7u) == 0u) {
    const uint64_t* wd = reinterpret_cast<const uint64_t*>(t.nflag + base);
    for (; p + 8 <= n; p += 8) s += __popcll((wd[p >> 3] >> 1) & 0x0101010101010101ull);  // F_SPECIAL = bit 1
  }
  for (; p < n; ++p) s += (t.nflag[base + p] & F_SPECIAL) != 0;
  return s;
}

// Inputs of node x for a plane kernel running as BC number `bc` of this step:
// nodes already rebuilt by an earlier BC of the same step hold plain inputs.
template <typename ST>
__device__ __forceinline__ void node_inputs(const ST* pop, const Grid& g, const Tables& t, int parity, int first,
                                            int64_t x, int bc, float f[19], uint32_t want = 0x7FFFFu) {
  const uint8_t fl = t.nflag[x];
  const int64_t nynz = (int64_t)g.ny * g.nz;
  const int i = (int)(x / nynz), j = (int)((x / g.nz) % g.ny), k = (int)(x % g.nz);
  const int fb = fl >> F_BC_SHIFT;
  const bool plain = fb != 0 && fb - 1 < bc;
  const bool special = (fl & F_SPECIAL) != 0;
  int4 m = make_int4(0, 0, 0, 0), b = make_int4(0, 0, 0, 0);
  if (special) {
    const int64_t s = special_index_scan(t, x);
    m = t.sp_mask[s];
    b = t.sp_base[s];
  }
  compose_inputs(pop, g, t, parity, first, i, j, k, special, plain, m, b, f, want);
}

template <typename ST>
__device__ __forceinline__ void store_inputs(ST* pop, const Grid& g, const Tables& t, int parity, int64_t x,
                                             const float f[19]) {
  const int64_t nynz = (int64_t)g.ny * g.nz;
  const int i = (int)(x / nynz), j = (int)((x / g.nz) % g.ny), k = (int)(x % g.nz);
#pragma unroll
  for (int q = 0; q < 19; ++q) store_plain(pop, g, t, parity, i, j, k, q, f[q]);
}

// Reference-layout copy of the populations entering the next step
// (solid nodes hold W, as after lbm_kernels.cu:597-609).
template <typename ST>
__global__ __launch_bounds__(kBlock) void aa_materialize_kernel(const ST* __restrict__ pop, Grid g, Tables t,
                                                                int parity, int first, float* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = x < g.N;
  const uint8_t fl = live ? t.nflag[x] : (uint8_t)F_SOLID;
  const bool special = (fl & F_SPECIAL) != 0;
  const uint64_t bal = __ballot(special);
  const int lane = threadIdx.x & 63;
  int4 m = make_int4(0, 0, 0, 0), b = make_int4(0, 0, 0, 0);
  if (special) {
    const int64_t sidx = t.wave_base[x >> 6] + __popcll(bal & ((1ull << lane) - 1ull));
    m = t.sp_mask[sidx];
    b = t.sp_base[sidx];
  }
  if (!live) return;
  if (fl & F_SOLID) {
#pragma unroll
    for (int q = 0; q < 19; ++q) out[q * g.N + x] = WQ(q);
    return;
  }
  const int64_t nynz = (int64_t)g.ny * g.nz;
  const int i = (int)(x / nynz), j = (int)((x / g.nz) % g.ny), k = (int)(x % g.nz);
  float f[19];
  compose_inputs(pop, g, t, parity, first, i, j, k, special, (fl >> F_BC_SHIFT) != 0, m, b, f);
#pragma unroll
  for (int q = 0; q < 19; ++q) out[q * g.N + x] = f[q];
}

// Load reference-layout populations as the input of an even step; link
// inputs go to A (used alone on the first step).
template <typename ST>
__global__ __launch_bounds__(kBlock) void aa_load_kernel(ST* __restrict__ pop, Grid g, Tables t,
                                                         const float* __restrict__ in) {
  const int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool live = x < g.N;
  const uint8_t fl = live ? t.nflag[x] : (uint8_t)F_SOLID;
  const bool special = (fl & F_SPECIAL) != 0;
  const uint64_t bal = __ballot(special);
  const int lane = threadIdx.x & 63;
  if (!live) return;
  const int64_t nynz = (int64_t)g.ny * g.nz;
  const int i = (int)(x / nynz), j = (int)((x / g.nz) % g.ny), k = (int)(x % g.nz);
  int64_t self;
  if (node_storage(g, t, i, j, k, 0, 0, 0, self)) {
#pragma unroll
    for (int q = 0; q < 19; ++q) Store<ST>::st(pop + (int64_t)q * g.Ns + self, in[q * g.N + x], q);
  }
  if (special) {
    const int64_t sidx = t.wave_base[x >> 6] + __popcll(bal & ((1ull << lane) - 1ull));
    const int4 m = t.sp_mask[sidx], b = t.sp_base[sidx];
    const uint32_t L = (uint32_t)m.x, V = (uint32_t)m.z;
    int64_t lp = b.x, lv = b.w;
    for (int q = 0; q < 19; ++q)
      if ((L >> q) & 1u) t.A[((V >> q) & 1u) ? lv++ : lp++] = in[q * g.N + x];
  }
}

// ---- inlet / outlet planes on the AA state ---------------------------------
__device__ __forceinline__ int64_t plane_flat(const Grid& g, int axis, int plane, int pa, int pb) {
  const int64_t nynz = (int64_t)g.ny * g.nz;
  if (axis == 0) return plane * nynz + (int64_t)pa * g.nz + pb;
  if (axis == 1) return pa * nynz + (int64_t)plane * g.nz + pb;
  return pa * nynz + (int64_t)pb * g.nz + plane;
}

__device__ __forceinline__ float qdiag(int c) { return c != 0 ? 0.66666667f : -0.33333333f; }

// Regularised Zou/He on the AA state (same node update as zou_he_kernel in
// lbm_ref.hip, lbm_kernels.cu:791-969).
template <typename ST>
__global__ __launch_bounds__(kPlaneBlock) void aa_zou_he_kernel(ST* __restrict__ pop, Grid g, Tables t, int parity,
                                                                int first, int bc, int axis, int plane, int interior,
                                                                int bc_type, float u_in, float rho_out, int sign,
                                                                int n1, int n2) {
  const int tid = blockIdx.x * kPlaneBlock + threadIdx.x;
  if (tid >= n1 * n2) return;
  const int64_t fb = plane_flat(g, axis, plane, tid / n2, tid % n2);
  if (t.nflag[fb] & F_SOLID) return;
  const int unknown = plane == 0 ? 1 : -1, known = -unknown;
  // the unknown inputs are rebuilt below before anything reads them: not loaded
  uint32_t rebuilt = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) rebuilt |= (C_AX(q, axis) == unknown ? 1u : 0u) << q;
  float fi[19];
  node_inputs(pop, g, t, parity, first, fb, bc, fi, 0x7FFFFu & ~rebuilt);
  float s0 = 0.f, sk = 0.f;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const int cq = C_AX(q, axis);
    if (cq == 0) s0 += fi[q];
    if (cq == known) sk += fi[q];
  }
  float rho_bc, uvec[3] = {0.f, 0.f, 0.f};
  if (bc_type == 0) {
    rho_bc = fminf(fmaxf((s0 + 2.f * sk) / (1.f - fabsf(u_in) + 1e-12f), 0.2f), 5.0f);
    uvec[axis] = u_in;
  } else {
    rho_bc = rho_out;
    const float uax = (float)known * ((s0 + 2.f * sk) / (rho_bc + 1e-12f) - 1.f);
    const int64_t fin = plane_flat(g, axis, interior, tid / n2, tid % n2);
    float gi[19];
    node_inputs(pop, g, t, parity, first, fin, bc, gi);
    if (uax * (float)sign < 0.f) {
      store_inputs(pop, g, t, parity, fb, gi);
      return;
    }
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
  float out[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const float s = qdiag(CX(q)) * pxx + qdiag(CY(q)) * pyy + qdiag(CZ(q)) * pzz +
                    2.f * ((float)(CX(q) * CY(q)) * pxy + (float)(CX(q) * CZ(q)) * pxz + (float)(CY(q) * CZ(q)) * pyz);
    out[q] = feq[q] + WQ(q) * s * 4.5f;
  }
  store_inputs(pop, g, t, parity, fb, out);
}

// Convective outlet on the AA state, pass 1 (copy + partial sums) and pass 2
// (mass correction); same semantics as convective_pass1/2 in lbm_ref.hip. A plane
// across z puts every input of every outlet node in its own cache line, so the copied
// inputs do not go back to their slots in pass 1: pass 1 writes them, as the storage
// type would return them, and their slot indices to a plane buffer ([19][plane] each,
// coalesced), and pass 2 scales them from there into the slots. The slots are then
// touched twice per step (pass 1 loads, pass 2 stores) instead of four times.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
template <typename ST>
__global__ __launch_bounds__(kPlaneBlock) void aa_convective_pass1(const ST* __restrict__ pop, Grid g, Tables t,
                                                                   int parity, int first, int bc, int axis,
                                                                   int outlet, int interior, int unknown_cax, int n1,
                                                                   int n2, double* __restrict__ partial,
                                                                   float* __restrict__ vals,
                                                                   uint32_t* __restrict__ slots) {
  const int tid = blockIdx.x * kPlaneBlock + threadIdx.x;
  const int P = n1 * n2;
  double r = 0.0, cnt = 0.0;
  if (tid < P) {
    const int64_t fo = plane_flat(g, axis, outlet, tid / n2, tid % n2);
    if (!(t.nflag[fo] & F_SOLID)) {
      const int64_t fi = plane_flat(g, axis, interior, tid / n2, tid % n2);
      // of the interior node only the five inputs copied to the outlet are read
      uint32_t unknown = 0;
#pragma unroll
      for (int q = 0; q < 19; ++q) unknown |= (C_AX(q, axis) == unknown_cax ? 1u : 0u) << q;
      float a[19], b2[19];
      node_inputs(pop, g, t, parity, first, fo, bc, a);
      node_inputs(pop, g, t, parity, first, fi, bc, b2, unknown);
      float rho = 0.f;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        if (C_AX(q, axis) == unknown_cax) a[q] = b2[q];
        rho += a[q];
      }
      const int64_t nynz = (int64_t)g.ny * g.nz;
      const int i = (int)(fo / nynz), j = (int)((fo / g.nz) % g.ny), k = (int)(fo % g.nz);
      const int d = parity == 0 ? 0 : -1;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        int64_t s;
        const bool has = node_storage(g, t, i, j, k, d * CX(q), d * CY(q), d * CZ(q), s);
        slots[q * P + tid] = has ? (uint32_t)s : kNoSlot;
        vals[q * P + tid] = Store<ST>::rt(a[q], q);
      }
      r = rho; cnt = 1.0;
    } else {
#pragma unroll
      for (int q = 0; q < 19; ++q) slots[q * P + tid] = kNoSlot;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    r += __shfl_down(r, off, 64);
    cnt += __shfl_down(cnt, off, 64);
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r;
    partial[2 * blockIdx.x + 1] = cnt;
  }
}

template <typename ST>
__global__ __launch_bounds__(kPlaneBlock) void aa_convective_pass2(ST* __restrict__ pop, Grid g, int parity, int n1,
                                                                   int n2, float rho_out,
                                                                   const double* __restrict__ partial, int nparts,
                                                                   const float* __restrict__ vals,
                                                                   const uint32_t* __restrict__ slots) {
  // every block folds the per-block partials the same way (lane l sums partials
  // l, l + 64, ..., then a fixed shuffle tree), so all blocks apply the same factor
  double a = 0.0, b = 0.0;
  for (int w = threadIdx.x; w < nparts; w += 64) { a += partial[2 * w]; b += partial[2 * w + 1]; }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  a = __shfl(a, 0, 64);
  b = __shfl(b, 0, 64);
  const float cnt = (float)b, sum = (float)a;
  if (cnt < 0.5f || sum < 1e-10f) return;
  const float corr = rho_out * cnt / sum;
  const int tid = blockIdx.x * kPlaneBlock + threadIdx.x;
  const int P = n1 * n2;
  if (tid >= P) return;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const uint32_t s = slots[q * P + tid];
    if (s == kNoSlot) continue;
    const int dir = parity == 0 ? q : OPP(q);
    Store<ST>::st(pop + (int64_t)dir * g.Ns + s, vals[q * P + tid] * corr, dir);
  }
}

// ---- host-side argument packing ----------------------------------------------
int make_grid(Grid& g, int nx, int ny, int nz, int px, int py, int pz) {
  FC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "aa: grid extents must be positive");
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.gx = px ? 0 : 1; g.gy = py ? 0 : 1; g.gz = pz ? 0 : 1;
  g.PY = ny + 2 * g.gy; g.PZ = nz + 2 * g.gz;
  g.Ns = (int64_t)(nx + 2 * g.gx) * g.PY * g.PZ;
  g.N = (int64_t)nx * ny * nz;
  return 0;
}

}  // namespace

// ===========================================================================
// C ABI. Storage type code: 0 FP32, 1 FP16 deviation, 2 BF16 deviation.
// ===========================================================================
extern "C" {

struct fc_aa_tables {
  const uint8_t* nflag;
  const int32_t* wave_base;
  const int32_t* wave_meta;  // [ceil(N/64) + 1][4]
  const int32_t* lk_code;
  const int32_t* s_code;
  const int32_t* f_code;
  const int32_t* lane_nodes;  // [n_lanes], or null when lane r is node r
  int64_t n_lanes;
  const int32_t* sp_mask;  // [n_special][4]
  const int32_t* sp_base;  // [n_special][4]
  const float* w1;
  const float* w2;
  float* A;
  float* B;
  const int32_t* ff_idx;
  int64_t n_v;
  const int32_t* row_meta;  // [rows * wpr][2], or null for dense storage
  int wpr;
  int64_t n_store;
};

struct fc_aa_collision {
  int kind;
  float om1, om2;
  float S[19];
  float cs_dx2;
  float force[3];
  int use_force;
  const float* sponge;
  int sponge_axis;
};

int64_t fc_aa_slot_count(int nx, int ny, int nz, int px, int py, int pz) {
  Grid g;
  if (make_grid(g, nx, ny, nz, px, py, pz)) return -1;
  return g.Ns;
}

static int pack(Grid& g, Tables& t, const fc_aa_tables* tb, int nx, int ny, int nz, int px, int py, int pz) {
  if (int rc = make_grid(g, nx, ny, nz, px, py, pz)) return rc;
  FC_REQUIRE(tb && tb->nflag && tb->wave_base && tb->wave_meta, "aa: node flag tables are required");
  t.nflag = tb->nflag; t.wave_base = tb->wave_base; t.wave_meta = (const int4*)tb->wave_meta;
  t.lk_code = tb->lk_code; t.s_code = tb->s_code; t.f_code = tb->f_code;
  t.lane_nodes = tb->lane_nodes; t.n_lanes = tb->n_lanes;
  FC_REQUIRE(t.n_lanes >= 0 && t.n_lanes <= g.N, "aa: lane count out of range");
  FC_REQUIRE(t.lane_nodes || t.n_lanes == g.N, "aa: lane list required unless there is one lane per node");
  t.sp_mask = (const int4*)tb->sp_mask; t.sp_base = (const int4*)tb->sp_base;
  t.w1 = tb->w1; t.w2 = tb->w2; t.A = tb->A; t.B = tb->B; t.ff_idx = tb->ff_idx; t.n_v = tb->n_v;
  t.row_meta = (const uint2*)tb->row_meta;
  t.wpr = tb->wpr;
  if (t.row_meta) {
    FC_REQUIRE(t.wpr == (g.PZ + 31) / 32, "aa: compact storage needs %d words per row (got %d)", (g.PZ + 31) / 32,
               t.wpr);
    FC_REQUIRE(tb->n_store > 0 && tb->n_store <= g.Ns, "aa: compact storage count out of range");
    g.Ns = tb->n_store;  // stride between the direction arrays
  }
  return 0;
}

int fc_aa_step(void* pop, int storage, int nx, int ny, int nz, int px, int py, int pz, const fc_aa_tables* tb,
               const fc_aa_collision* col, int parity, int first, float* rho, float* u, float* v, float* w,
               int solid_macros, hipStream_t stream) {
  Grid g; Tables t;
  if (int rc = pack(g, t, tb, nx, ny, nz, px, py, pz)) return rc;
  FC_REQUIRE(pop, "aa_step: null population buffer");
  FC_REQUIRE(g.Ns < ((int64_t)1 << 30), "aa_step: padded grid of %lld nodes exceeds 2^30", (long long)g.Ns);
  FC_REQUIRE((int64_t)g.PY * g.PZ < ((int64_t)1 << 24) && nx + 2 < (1 << 12),
             "aa_step: padded extents exceed the 24-bit slot arithmetic");
  FC_REQUIRE(col && col->kind >= 0 && col->kind <= 3, "aa_step: collision kind must be 0..3");
  FC_REQUIRE(parity == 0 || parity == 1, "aa_step: parity must be 0 or 1");
  FC_REQUIRE((rho && u && v && w) || (!rho && !u && !v && !w), "aa_step: macro outputs must be all set or all null");
  Collision c;
  c.kind = col->kind; c.om1 = col->om1; c.om2 = col->om2;
  for (int k = 0; k < 19; ++k) c.S[k] = col->S[k];
  c.cs_dx2 = col->cs_dx2;
  for (int d = 0; d < 3; ++d) c.F[d] = col->force[d];
  c.use_force = col->use_force; c.sponge = col->sponge; c.sponge_axis = col->sponge_axis;
  // with one lane per node the step itself writes the solid nodes' macros
  if (rho && solid_macros && t.lane_nodes) {
    hipLaunchKernelGGL(aa_solid_macros_kernel, dim3(grid_for(g.N, kBlock)), dim3(kBlock), 0, stream, t.nflag, g.N,
                       rho, u, v, w);
    if (int rc = check_launch("aa_step/solid macros")) return rc;
  }
  if (t.n_lanes == 0) return 0;
  const unsigned grid = grid_for(t.n_lanes, kBlock);
#define FC_STEP(T, K)                                                                                            \
  if (t.row_meta)                                                                                                  \
    hipLaunchKernelGGL((aa_step_kernel<T, K, true>), dim3(grid), dim3(kBlock), 0, stream, (T*)pop, g, t, c, parity, \
                       first, rho, u, v, w);                                                                       \
  else                                                                                                             \
    hipLaunchKernelGGL((aa_step_kernel<T, K, false>), dim3(grid), dim3(kBlock), 0, stream, (T*)pop, g, t, c,        \
                       parity, first, rho, u, v, w)
#define FC_STEP_KIND(T)                      \
  switch (c.kind) {                          \
    case 0: FC_STEP(T, 0); break;            \
    case 1: FC_STEP(T, 1); break;            \
    case 2: FC_STEP(T, 2); break;            \
    default: FC_STEP(T, 3); break;           \
  }
  switch (storage) {
    case 0: FC_STEP_KIND(float); break;
    case 1: FC_STEP_KIND(__half); break;
    case 2: FC_STEP_KIND(__hip_bfloat16); break;
    default: return fc::set_error(1, "aa_step: storage code must be 0 (fp32), 1 (fp16) or 2 (bf16)");
  }
#undef FC_STEP_KIND
#undef FC_STEP
  return check_launch("aa_step");
}

int fc_aa_materialize(const void* pop, int storage, int nx, int ny, int nz, int px, int py, int pz,
                      const fc_aa_tables* tb, int parity, int first, float* out, hipStream_t stream) {
  Grid g; Tables t;
  if (int rc = pack(g, t, tb, nx, ny, nz, px, py, pz)) return rc;
  FC_REQUIRE(pop && out, "aa_materialize: null pointer argument");
  const unsigned grid = grid_for(g.N, kBlock);
  switch (storage) {
    case 0: hipLaunchKernelGGL(aa_materialize_kernel<float>, dim3(grid), dim3(kBlock), 0, stream, (const float*)pop,
                               g, t, parity, first, out); break;
    case 1: hipLaunchKernelGGL(aa_materialize_kernel<__half>, dim3(grid), dim3(kBlock), 0, stream,
                               (const __half*)pop, g, t, parity, first, out); break;
    case 2: hipLaunchKernelGGL(aa_materialize_kernel<__hip_bfloat16>, dim3(grid), dim3(kBlock), 0, stream,
                               (const __hip_bfloat16*)pop, g, t, parity, first, out); break;
    default: return fc::set_error(1, "aa_materialize: bad storage code");
  }
  return check_launch("aa_materialize");
}

int fc_aa_load(void* pop, int storage, int nx, int ny, int nz, int px, int py, int pz, const fc_aa_tables* tb,
               const float* in, hipStream_t stream) {
  Grid g; Tables t;
  if (int rc = pack(g, t, tb, nx, ny, nz, px, py, pz)) return rc;
  FC_REQUIRE(pop && in, "aa_load: null pointer argument");
  const unsigned grid = grid_for(g.N, kBlock);
  switch (storage) {
    case 0: hipLaunchKernelGGL(aa_load_kernel<float>, dim3(grid), dim3(kBlock), 0, stream, (float*)pop, g, t, in);
      break;
    case 1: hipLaunchKernelGGL(aa_load_kernel<__half>, dim3(grid), dim3(kBlock), 0, stream, (__half*)pop, g, t, in);
      break;
    case 2: hipLaunchKernelGGL(aa_load_kernel<__hip_bfloat16>, dim3(grid), dim3(kBlock), 0, stream,
                               (__hip_bfloat16*)pop, g, t, in); break;
    default: return fc::set_error(1, "aa_load: bad storage code");
  }
  return check_launch("aa_load");
}

int fc_aa_zou_he(void* pop, int storage, int nx, int ny, int nz, int px, int py, int pz, const fc_aa_tables* tb,
                 int parity, int first, int bc_order, int axis, int plane_idx, int interior_idx, int bc_type,
                 float u_in_lat, float rho_out_lat, int sign, hipStream_t stream) {
  Grid g; Tables t;
  if (int rc = pack(g, t, tb, nx, ny, nz, px, py, pz)) return rc;
  FC_REQUIRE(axis >= 0 && axis <= 2 && (bc_type == 0 || bc_type == 1), "aa_zou_he: bad axis or bc_type");
  const int n_ax = axis == 0 ? nx : (axis == 1 ? ny : nz);
  FC_REQUIRE(plane_idx >= 0 && plane_idx < n_ax && interior_idx >= 0 && interior_idx < n_ax,
             "aa_zou_he: plane index out of range");
  const int n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
  const unsigned grid = grid_for((int64_t)n1 * n2, kPlaneBlock);
#define FC_ZH(T)                                                                                                  \
  hipLaunchKernelGGL(aa_zou_he_kernel<T>, dim3(grid), dim3(kPlaneBlock), 0, stream, (T*)pop, g, t, parity, first,      \
                     bc_order, axis, plane_idx, interior_idx, bc_type, u_in_lat, rho_out_lat, sign, n1, n2)
  switch (storage) {
    case 0: FC_ZH(float); break;
    case 1: FC_ZH(__half); break;
    case 2: FC_ZH(__hip_bfloat16); break;
    default: return fc::set_error(1, "aa_zou_he: bad storage code");
  }
#undef FC_ZH
  return check_launch("aa_zou_he");
}

// [2 per block] FP64 partial sums, then the plane buffer: [19][n1 n2] FP32 values and
// [19][n1 n2] slot indices
size_t fc_aa_convective_workspace_bytes(int n1, int n2) {
  if (n1 <= 0 || n2 <= 0) return 0;
  const size_t P = (size_t)n1 * (size_t)n2;
  return sizeof(double) * 2 * (size_t)grid_for((int64_t)P, kPlaneBlock) + 19 * P * (sizeof(float) + sizeof(uint32_t));
}

int fc_aa_convective(void* pop, int storage, int nx, int ny, int nz, int px, int py, int pz,
                     const fc_aa_tables* tb, int parity, int first, int bc_order, int axis, int outlet_idx,
                     int interior_idx, int unknown_cax, float rho_out_lat, void* workspace, size_t ws_bytes,
                     hipStream_t stream) {
  Grid g; Tables t;
  if (int rc = pack(g, t, tb, nx, ny, nz, px, py, pz)) return rc;
  FC_REQUIRE(axis >= 0 && axis <= 2 && (unknown_cax == 1 || unknown_cax == -1), "aa_convective: bad axis/unknown_cax");
  const int n_ax = axis == 0 ? nx : (axis == 1 ? ny : nz);
  FC_REQUIRE(outlet_idx >= 0 && outlet_idx < n_ax && interior_idx >= 0 && interior_idx < n_ax,
             "aa_convective: plane index out of range");
  const int n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
  const size_t need = fc_aa_convective_workspace_bytes(n1, n2);
  FC_REQUIRE(workspace && ws_bytes >= need, "aa_convective: workspace of %zu bytes required", need);
  FC_REQUIRE(g.Ns < (int64_t)kNoSlot, "aa_convective: more than 2^32 - 2 slots per direction");
  const unsigned grid = grid_for((int64_t)n1 * n2, kPlaneBlock);
  double* partial = (double*)workspace;
  float* vals = (float*)(partial + 2 * (size_t)grid);
  uint32_t* slots = (uint32_t*)(vals + 19 * (size_t)n1 * n2);
#define FC_CV(T)                                                                                                 \
  hipLaunchKernelGGL(aa_convective_pass1<T>, dim3(grid), dim3(kPlaneBlock), 0, stream, (const T*)pop, g, t, parity,  \
                     first, bc_order, axis, outlet_idx, interior_idx, unknown_cax, n1, n2, partial, vals, slots);  \
  if (int rc = check_launch("aa_convective/pass1")) return rc;                                                    \
  hipLaunchKernelGGL(aa_convective_pass2<T>, dim3(grid), dim3(kPlaneBlock), 0, stream, (T*)pop, g, parity, n1, n2,    \
                     rho_out_lat, (const double*)partial, (int)grid, (const float*)vals, (const uint32_t*)slots)
  switch (storage) {
    case 0: FC_CV(float); break;
    case 1: FC_CV(__half); break;
    case 2: FC_CV(__hip_bfloat16); break;
    default: return fc::set_error(1, "aa_convective: bad storage code");
  }
#undef FC_CV
  return check_launch("aa_convective/pass2");
}

}  // extern "C"
