// unet.hip — element and resampling kernels around the convolutions of the
// learned preconditioner (channels-last BF16 activations).
//
//  pack_input:  x[p] = {r[p] sqrt(n) / ||r||, feat1[p], mask[p], 0, ...} (8 channels):
//               the residual at unit RMS, so that in BF16 it is not swamped by the
//               O(1) feature channels; ||r||^2 is a device FP64 scalar (no host round trip).
//  avgpool2:    2x2x2 mean (FP32 sum of 8 BF16 values, one rounding).
//  scale:       out[p] = r[p] * d[p] (symmetric diagonal scaling D^-1/2 r).
//  unpack:      z[p] = s * y[p, 0] * mask[p] * d[p], s = sqrt(scal[slot] / n);
//  unpack_add:  the same added to z (a V-cycle's output plus the network's correction);
//  residual_scale: (r - A z) * mask * d, the residual a V-cycle's output z leaves.
#include "fc_common.hpp"

using namespace fc;

namespace {

constexpr int kBlock = 256;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float norm_from(const double* scal, int slot) {
  const double s = scal[slot];
  return s > 0.0 ? (float)sqrt(s) : 0.f;
}

__global__ __launch_bounds__(kBlock) void pack_input_kernel(const float* __restrict__ r,
                                                            const float* __restrict__ feat1,
                                                            const float* __restrict__ mask,
                                                            const double* __restrict__ scal, int slot, int64_t n,
                                                            __bf16* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const float nr = norm_from(scal, slot);
  const float inv = nr > 0.f ? sqrtf((float)n) / nr : 0.f;
  bf16x8 v;
  v[0] = (__bf16)(r[p] * inv);
  v[1] = (__bf16)(feat1 ? feat1[p] : 0.f);
  v[2] = (__bf16)(mask ? mask[p] : 1.f);
#pragma unroll
  for (int j = 3; j < 8; ++j) v[j] = (__bf16)0.f;
  *(bf16x8*)(out + p * 8) = v;
}

// one thread per (coarse position, 8-channel group)
__global__ __launch_bounds__(kBlock) void avgpool2_kernel(const __bf16* __restrict__ in, int B, int D, int H, int W,
                                                          int C, __bf16* __restrict__ out) {
  const int D2 = D >> 1, H2 = H >> 1, W2 = W >> 1, G = C >> 3;
  const int64_t n = (int64_t)B * D2 * H2 * W2 * G;
  const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (id >= n) return;
  const int g = (int)(id % G);
  int64_t r = id / G;
  const int w = (int)(r % W2); r /= W2;
  const int h = (int)(r % H2); r /= H2;
  const int d = (int)(r % D2);
  const int b = (int)(r / D2);
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int dd = 2 * d + (o >> 2), hh = 2 * h + ((o >> 1) & 1), ww = 2 * w + (o & 1);
    const int64_t p = (((int64_t)b * D + dd) * H + hh) * W + ww;
    const bf16x8 v = *(const bf16x8*)(in + p * C + g * 8);
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] += (float)v[j];
  }
  bf16x8 o8;
#pragma unroll
  for (int j = 0; j < 8; ++j) o8[j] = (__bf16)(acc[j] * 0.125f);
  *(bf16x8*)(out + id * 8) = o8;
}

__global__ __launch_bounds__(kBlock) void unpack_kernel(const float* __restrict__ y, int cy,
                                                        const float* __restrict__ mask,
                                                        const float* __restrict__ d,
                                                        const double* __restrict__ scal, int slot, int64_t n,
                                                        float* __restrict__ z) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const float nr = norm_from(scal, slot) / sqrtf((float)n);
  const float m = mask ? mask[p] : 1.f;
  const float dd = d ? d[p] : 1.f;
  z[p] = nr * y[p * cy] * m * dd;
}

// z[p] += s y[p, 0] mask[p] d[p]: the network's correction added to a V-cycle's output
__global__ __launch_bounds__(kBlock) void unpack_add_kernel(const float* __restrict__ y, int cy,
                                                            const float* __restrict__ mask,
                                                            const float* __restrict__ d,
                                                            const double* __restrict__ scal, int slot, int64_t n,
                                                            float* __restrict__ z) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const float nr = norm_from(scal, slot) / sqrtf((float)n);
  const float m = mask ? mask[p] : 1.f;
  const float dd = d ? d[p] : 1.f;
  z[p] = z[p] + nr * y[p * cy] * m * dd;
}

// out[p] = (r[p] - Lo[p]) mask[p] d[p]: the residual left by a V-cycle's output z (Lo = A z)
// in the symmetrically scaled variables
__global__ __launch_bounds__(kBlock) void residual_scale_kernel(const float* __restrict__ r,
                                                                const float* __restrict__ Lo,
                                                                const float* __restrict__ mask,
                                                                const float* __restrict__ d, int64_t n,
                                                                float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  const float m = mask ? mask[p] : 1.f;
  out[p] = (r[p] - Lo[p]) * m * d[p];
}

__global__ __launch_bounds__(kBlock) void scale_kernel(const float* __restrict__ r, const float* __restrict__ d,
                                                       int64_t n, float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= n) return;
  out[p] = r[p] * d[p];
}

}  // namespace

extern "C" {

int fc_unet_pack_input(const float* r, const float* feat1, const float* mask, const double* scal, int slot,
                       int64_t n, void* out, hipStream_t stream) {
  FC_REQUIRE(r && scal && out && n >= 0 && slot >= 0, "unet_pack_input: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(pack_input_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, r, feat1, mask, scal,
                     slot, n, (__bf16*)out);
  return check_launch("unet_pack_input");
}

int fc_unet_avgpool2(const void* in, int B, int D, int H, int W, int C, void* out, hipStream_t stream) {
  FC_REQUIRE(in && out, "unet_avgpool2: null pointer argument");
  FC_REQUIRE(B > 0 && D > 1 && H > 1 && W > 1 && D % 2 == 0 && H % 2 == 0 && W % 2 == 0,
             "unet_avgpool2: extents must be even and >= 2");
  FC_REQUIRE(C > 0 && C % 8 == 0, "unet_avgpool2: channels must be a multiple of 8");
  const int64_t n = (int64_t)B * (D / 2) * (H / 2) * (W / 2) * (C / 8);
  hipLaunchKernelGGL(avgpool2_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, (const __bf16*)in, B, D,
                     H, W, C, (__bf16*)out);
  return check_launch("unet_avgpool2");
}

int fc_unet_unpack(const float* y, int cy, const float* mask, const float* d, const double* scal, int slot,
                   int64_t n, float* z, hipStream_t stream) {
  FC_REQUIRE(y && scal && z && cy > 0 && n >= 0 && slot >= 0, "unet_unpack: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(unpack_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, y, cy, mask, d, scal, slot,
                     n, z);
  return check_launch("unet_unpack");
}

int fc_unet_unpack_add(const float* y, int cy, const float* mask, const float* d, const double* scal, int slot,
                       int64_t n, float* z, hipStream_t stream) {
  FC_REQUIRE(y && scal && z && cy > 0 && n >= 0 && slot >= 0, "unet_unpack_add: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(unpack_add_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, y, cy, mask, d, scal,
                     slot, n, z);
  return check_launch("unet_unpack_add");
}

int fc_unet_residual_scale(const float* r, const float* Lo, const float* mask, const float* d, int64_t n, float* out,
                           hipStream_t stream) {
  FC_REQUIRE(r && Lo && d && out && n >= 0, "unet_residual_scale: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(residual_scale_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, r, Lo, mask, d, n,
                     out);
  return check_launch("unet_residual_scale");
}

int fc_unet_scale(const float* r, const float* d, int64_t n, float* out, hipStream_t stream) {
  FC_REQUIRE(r && d && out && n >= 0, "unet_scale: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, r, d, n, out);
  return check_launch("unet_scale");
}

}  // extern "C"
