// mlp_kernels.h -- layout and device helpers of the canopy surrogate MLP
// (24 -> 128 -> 128 -> 128 -> 8, SiLU), shared by the training kernels
// (mlp.hip) and the inference the hybrid step runs (cbm_kernel.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbm_mlp {

constexpr int NIN = 24, HID = 128, NOUT = 8;
constexpr int LAYER_IN[4] = {NIN, HID, HID, HID};
constexpr int LAYER_OUT[4] = {HID, HID, HID, NOUT};
// packed parameter vector: W1 (24x128), b1, W2 (128x128), b2, W3, b3,
// W4 (128x8), b4; each W is row-major [in][out]
constexpr int LAYER_W[4] = {0, NIN * HID + HID, 2 * (NIN * HID + HID) - NIN * HID + HID * HID,
                            NIN * HID + HID + 2 * (HID * HID + HID)};
constexpr int LAYER_B[4] = {LAYER_W[0] + NIN * HID, LAYER_W[1] + HID * HID, LAYER_W[2] + HID * HID,
                            LAYER_W[3] + HID * NOUT};
constexpr int NPARAM = LAYER_B[3] + NOUT;
static_assert(NPARAM == 37256, "24-128-128-128-8 has 37,256 parameters");
static_assert(LAYER_W[2] == LAYER_B[1] + HID, "packed layout");
// normalisation vector: mu_x[24], sd_x[24], mu_y[8], sd_y[8]
constexpr int NNORM = 2 * NIN + 2 * NOUT;

typedef float v16f __attribute__((ext_vector_type(16)));
typedef short v4s __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__device__ inline float to_f32(float x) { return x; }
__device__ inline float to_f32(uint16_t x) { return bf16_to_f32(x); }

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

}  // namespace cbm_mlp
