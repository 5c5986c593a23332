// mlp_kernels.h -- layout and device helpers of the canopy surrogate MLP
// (24 -> 128 -> 128 -> 128 -> 8, hard-swish), shared by the training kernels
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

// The training order: an endless sequence of epochs, each a permutation of
// the n samples chosen by (seed, epoch).  The permutation is a keyed
// bijection of [0, 2^bits) (2^bits the power of two at or above n) -- four
// rounds of multiply by an odd key, xor-shift right and add a key, each
// invertible modulo 2^bits -- restricted to [0, n) by cycle walking: an
// image at or above n is mapped again, and the walk ends inside [0, n)
// because it follows a cycle through the starting point.  Each step leaves
// [0, n) with probability at most 1/2, so the cap of 4096 steps is never
// reached in practice; it only keeps the loop finite.  The keys come from
// the splitmix64 finaliser (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__host__ __device__ inline int64_t shuffled_sample(uint64_t seed, int64_t pos, int64_t n) {
  const int64_t epoch = pos / n, i = pos - epoch * n;
  int bits = 1;
  while (bits < 62 && (int64_t(1) << bits) < n) bits++;
  const uint64_t mask = (uint64_t(1) << bits) - 1;
  const int sh = (bits + 1) / 2;
  uint64_t key[4];
  for (int r = 0; r < 4; r++) key[r] = mix64(seed ^ mix64((uint64_t)epoch * 4 + r));
  uint64_t x = (uint64_t)i;
  for (int step = 0; step < 4096; step++) {
    for (int r = 0; r < 4; r++) {
      x = (x * (key[r] | 1)) & mask;
      x ^= x >> sh;
      x = (x + (key[r] >> 32)) & mask;
    }
    if (x < (uint64_t)n) return (int64_t)x;
  }
  return i;
}

// The hidden layers' activation, hard-swish: h(z) = z min(max(z + 3, 0), 6) / 6,
// the piecewise-quadratic stand-in for SiLU (Howard et al., "Searching for
// MobileNetV3", ICCV 2019).  It costs four plain VALU instructions where
// SiLU costs an exponential and a reciprocal, which took half the inference
// kernel's VALU cycles.  h'(z) = 0 below -3, (2z + 3) / 6 between, 1 above.
// Written as z clamp(z/6 + 1/2, 0, 1), the same function: the clamp to [0, 1]
// folds into the fused multiply-add's output modifier, so h is two VALU
// instructions (the form z min(max(z + 3, 0), 6) / 6 takes four).
__host__ __device__ inline float hswish(float z) {
  return z * fminf(fmaxf(fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f), 1.0f);
}
__host__ __device__ inline float hswish_grad(float z) {
  return z < -3.0f ? 0.0f : z > 3.0f ? 1.0f : (2.0f * z + 3.0f) * (1.0f / 6.0f);
}

// a workgroup barrier that orders LDS only: __syncthreads() would also wait
// for every global load and store in flight (weights or inputs still
// arriving, results still being written)
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace cbm_mlp
