// mlp.hip -- the canopy surrogate: a 24 -> 128 -> 128 -> 128 -> 8 MLP with
// hard-swish, trained on the GPU: BF16 operands on the matrix cores
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation, FP32 master weights and
// Adam.  All kernels are hand-written for gfx950 wave64.
//
// A training batch takes two launches.
//
// rows_kernel: a workgroup of four waves takes the batch 32 rows at a time.
// It gathers and normalises their samples, runs the forward pass and the
// backward pass down to the first layer's pre-activation gradient, and forms
// its rows' share of every parameter's gradient.  Wave w owns features
// 32w .. 32w+31 of every 128-wide layer in both passes, so the activation's
// derivative it needs on the way back is the one it computed on the way
// forward, still in its registers; activations and gradients pass between
// layers through LDS (one workgroup barrier per layer).  Each wave loads its
// slices of the BF16 weights into registers once per launch.  The layer
// inputs H and the pre-activation gradients dZ of the 32 rows are also kept
// in LDS feature-major, and the workgroup forms dW = H^T dZ over its rows on
// the matrix cores and the bias gradients as FP32 sums of dZ, into its own
// slot of a partial-gradient buffer (a workgroup that takes several blocks
// of rows adds them up there).
//
// update_kernel: four threads per parameter sum the workgroups' partials in
// a fixed order and take the Adam step (or store the gradient), then
// refresh the BF16 copies the next batch loads.  Nothing is summed by
// atomics: a batch's gradient, loss and update are the same on every run.
//
// At 2,048 rows a batch runs 64 workgroups and is bound by latency, not by
// the matrix cores: the rows pass is a chain of seven dependent products,
// their epilogues and barriers, and the update waits on loads of partials
// written by every XCD.
//
// A product is computed transposed, D = W^T X^T (W [in][out], X [row][in]):
// lane l supplies A[l%32][k0 + 8(l/32) .. +7] and B[k0 + 8(l/32) .. +7][l%32],
// eight consecutive BF16 each, and receives D[8g + 4(l/32) + q][l%32] in
// register 4g + q, so a lane holds four consecutive features of one row (row
// l%32) per register group.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cable_mlp.h"
#include "mlp_kernels.h"

using namespace cbm_mlp;

namespace {

thread_local std::string g_err;
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define HIPCHECK(expr)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(CABLE_MLP_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// kernels

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

constexpr int RW = 32;                 // rows per block
constexpr int NWAVE = 4;               // waves of rows_kernel, one per 32 features
constexpr int AP = HID + 8;            // row-major LDS pitch (BF16): 68 dwords, conflict-free 16-byte reads
constexpr int TP = RW + 8;             // feature-major LDS pitch (BF16): 20 dwords, likewise
constexpr int K1 = 32;                 // layer 1's K (24 inputs) padded to a multiple of 16
constexpr int N4B = 16;                // layer 4's 8 outputs padded, as the K of the backward product
// BF16 working copies, zero-padded.  wT: W^T [out][in] of every layer for
// the forward pass -- W1^T [128][32], W2^T and W3^T [128][128], W4^T
// [32][128] with rows 8..31 zero.  wB: W [in][out] of layers 4, 3, 2 for the
// backward pass -- W4 [128][16] with columns 8..15 zero, W3 and W2 [128][128].
constexpr int WT_OFF[4] = {0, HID * K1, HID * K1 + HID * HID, HID * K1 + 2 * HID * HID};
constexpr int WT_SIZE = WT_OFF[3] + 32 * HID;
constexpr int WB_OFF[4] = {-1, HID * N4B + HID * HID, HID * N4B, 0};
constexpr int WB_SIZE = HID * N4B + 2 * HID * HID;
// feature-major LDS rows of a block: the inputs of layers 1..4 (layer 1's
// padded to 32) and the pre-activation gradients of layers 1..4 (layer 4's
// padded to 32); the pads are zero
constexpr int HT_OFF[4] = {0, K1, K1 + HID, K1 + 2 * HID};
constexpr int HT_ROWS = K1 + 3 * HID;
constexpr int DT_OFF[4] = {0, HID, 2 * HID, 3 * HID};
constexpr int DT_ROWS = 3 * HID + 32;
// 32x32 blocks of the weight gradients: layer 1 (one input block, 24 of its
// 32 rows real) x 4, layers 2 and 3 4 x 4, layer 4 4 x (one output block, 8
// of its 32 columns real)
constexpr int NBLOCKS = 4 + 16 + 16 + 4;
constexpr int NBIAS = 3 * HID + NOUT;
// a workgroup's slot of partial gradients, in the order the matrix cores
// leave them: the 40 weight-gradient blocks, each as the 16 registers of its
// 64 lanes (so a register's store is 256 contiguous bytes and a block's 4 KB;
// the pad rows and columns are stored too and never read), then the biases
// in layer order, then the sum of squared errors; padded to 256 bytes.
// (Stored in the parameter vector's own layout, each register's lanes write
// two 128-byte pieces 512 bytes apart and the stores took twice as long as
// the rest of the pass.)
constexpr int PBLK = 16 * 64;
constexpr int PB_OFF = NBLOCKS * PBLK;
constexpr int PLOSS = PB_OFF + NBIAS;
constexpr int PSTRIDE = (PLOSS + 1 + 63) / 64 * 64;
constexpr int GMAX = 256;              // most workgroups of rows_kernel: one per CU

struct RowsArgs {
  const float *xs, *ys;      // raw samples gathered through idx (training), or
  const int64_t *idx;        // through the shuffled order (shuffled) ...
  int64_t nsamples, pos0;    // ... positions pos0 + row of it
  uint64_t seed;
  int shuffled;
  const float *x, *y;        // ... or normalised rows (idx null)
  const float *norm;
  const float *w;            // FP32 parameters (the biases are read here)
  const uint16_t *wT, *wB;   // BF16 copies
  float *out;                // forward only: outputs [n][8]
  float *part;               // with the backward pass: partial gradients [gridDim.x][PSTRIDE]
  int n, backward;
};

struct UpdateArgs {
  const float *part;
  int npart, n;
  float *w, *m, *v, *g;      // parameters, Adam's moments, gradient
  uint16_t *wT, *wB;
  float *loss;               // the batch's mean-squared error (optional)
  int adam;                  // 1: Adam step at rate lr; 0: store the gradient in g
  float lr, b1, b2, eps, c1, c2;
};

// the sum, over the 32 lanes of a half-wave, of each of a lane's NV values.
// Each exchange step halves the values a lane carries: the lane keeps one
// half, adds its partner's copy of it and passes on the other half; after
// log2(NV) steps the rest is a plain butterfly.  The lane returns the total
// of value number vi; lanes that differ only in their low 5 - log2(NV) bits
// return the same total.  NV - 1 + 5 - log2(NV) exchanges in all.
template <int NV>
struct floats {
  typedef float type __attribute__((ext_vector_type(NV)));
};

// (the empty asm makes the two values opaque: without it the compiler folds
// the selects into a run-time index into v, a chain of compares over every
// element)
template <int NV>
__device__ __forceinline__ float half_sums(typename floats<NV>::type v, int lane, int &vi) {
  vi = 0;
  int m = 16;
#pragma unroll
  for (int h = NV / 2; h >= 1; h /= 2, m >>= 1) {
    const bool up = (lane & m) != 0;
#pragma unroll
    for (int t = 0; t < h; t++) {
      float lo = v[t], hi = v[t + h];
      asm("" : "+v"(lo), "+v"(hi));
      const float send = up ? lo : hi;
      const float keep = up ? hi : lo;
      v[t] = keep + __shfl_xor(send, m, 64);
    }
    vi += up ? h : 0;
  }
  float s = v[0];
#pragma unroll
  for (; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  return s;
}

// the block's contribution to partial gradient k of this workgroup (the
// first block it takes stores, later ones add; always the same thread)
__device__ __forceinline__ void part_add(const RowsArgs &p, int k, float v, bool first) {
  float *q = p.part + (int64_t)blockIdx.x * PSTRIDE + k;
  *q = first ? v : *q + v;
}

__device__ __forceinline__ v16f zero16() {
  v16f a;
#pragma unroll
  for (int r = 0; r < 16; r++) a[r] = 0.0f;
  return a;
}

// D = W^T X^T for this wave's 32 output features over the block's 32 rows;
// KS steps of K 16.  Two accumulators halve the dependent chain.
template <int KS>
__device__ __forceinline__ v16f product(const bf8 *W, const __bf16 *in, int col, int kh, v16f a0 = zero16()) {
  v16f a1 = zero16();
  // every operand read issued before the first product: left to itself the
  // scheduler reads each one just before its product and, at one wave per
  // SIMD, waits out every LDS latency
  bf8 xv[KS];
#pragma unroll
  for (int s = 0; s < KS; s++) xv[s] = *(const bf8 *)&in[col * AP + 16 * s + kh];
#pragma unroll
  for (int s = 0; s < KS; s++) {
    if (s & 1)
      a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[s], xv[s], a1, 0, 0, 0);
    else
      a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[s], xv[s], a0, 0, 0, 0);
  }
  __builtin_amdgcn_sched_group_barrier(0x0100, KS, 0);  // LDS reads
  __builtin_amdgcn_sched_group_barrier(0x0008, KS, 0);  // matrix products
  if (KS > 1) a0 += a1;
  return a0;
}

// a hidden layer forward: z = XW + b, h = hswish(z) into `out` (and into the
// block's feature-major H rows `ht` when the backward pass follows), with
// hswish'(z) kept in sg
template <int KS>
__device__ __forceinline__ void forward_layer(const bf8 *W, const __bf16 *in, __bf16 *out, const float *b,
                                              float *sg, __bf16 *ht, int wave, int col, int kh, int kq) {
  // the products add onto the bias
  v16f b16;
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const float4 bv = *(const float4 *)&b[32 * wave + 8 * g + kq];
    b16[4 * g + 0] = bv.x;
    b16[4 * g + 1] = bv.y;
    b16[4 * g + 2] = bv.z;
    b16[4 * g + 3] = bv.w;
  }
  const v16f acc = product<KS>(W, in, col, kh, b16);
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int j = 32 * wave + 8 * g + kq;
    bf4 hv;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float z = acc[4 * g + q];
      sg[4 * g + q] = hswish_grad(z);
      hv[q] = (__bf16)hswish(z);
    }
    *(bf4 *)&out[col * AP + j] = hv;
    if (ht)
#pragma unroll
      for (int q = 0; q < 4; q++) ht[(j + q) * TP + col] = hv[q];
  }
}

// a hidden layer backward: dZ = (dZ_next W^T) hswish'(z), into `out` (when a
// product follows) and the block's feature-major dZ rows `dt`; the bias
// gradient's FP32 share from this block into the partials at bias0
template <int KS>
__device__ __forceinline__ void backward_layer(const RowsArgs &p, const bf8 *W, const __bf16 *in, __bf16 *out,
                                               const float *sg, __bf16 *dt, int bias0, bool first, int wave,
                                               int lane, int col, int kh, int kq) {
  const v16f acc = product<KS>(W, in, col, kh);
  v16f dz;
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int i = 32 * wave + 8 * g + kq;
    bf4 dv;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      dz[4 * g + q] = acc[4 * g + q] * sg[4 * g + q];
      dv[q] = (__bf16)dz[4 * g + q];
    }
    if (out) *(bf4 *)&out[col * AP + i] = dv;
#pragma unroll
    for (int q = 0; q < 4; q++) dt[(i + q) * TP + col] = dv[q];
  }
  int vi;
  const float s = half_sums<16>(dz, lane, vi);
  if ((lane & 1) == 0) part_add(p, bias0 + 32 * wave + 8 * (vi >> 2) + kq + (vi & 3), s, first);
}

struct RowsLDS {
  __attribute__((aligned(16))) __bf16 act[2][RW * AP];
  __attribute__((aligned(16))) __bf16 Ht[HT_ROWS * TP];
  __attribute__((aligned(16))) __bf16 Dt[DT_ROWS * TP];
  __attribute__((aligned(16))) float bias[NBIAS];
  float tgt[RW * NOUT];
};

// this wave's share of layer L's weight gradient over the block's rows,
// dW = H^T dZ with K = the 32 rows, into the workgroup's slot of partials:
// the 32x32 blocks (ib, jb) with ib = wave (layers 2-4: input block) or
// jb = wave (layer 1: output block).  A later block's loads all go before
// its stores.  The stores are non-temporal: the next kernel reads them on
// every XCD, so nothing gains from keeping them in this XCD's L2 (measured
// 3 us less per pass at 64 workgroups).
template <int L>
__device__ __forceinline__ void weight_grad_blocks(const RowsArgs &p, const __bf16 *Ht, const __bf16 *Dt, int wave,
                                                   bool first, int lane, int col, int kh) {
  constexpr int NJ = L == 0 ? 1 : L == 3 ? 1 : 4;      // blocks per wave
  constexpr int B0 = L == 0 ? 0 : L == 1 ? 4 : L == 2 ? 20 : 36;
#pragma unroll
  for (int t = 0; t < NJ; t++) {
    const int ib = L == 0 ? 0 : wave, jb = L == 0 ? wave : L == 3 ? 0 : t;
    const int bk = B0 + (L == 0 ? wave : L == 3 ? wave : 4 * wave + t);
    v16f acc = zero16();
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const bf8 av = *(const bf8 *)&Ht[(HT_OFF[L] + 32 * ib + col) * TP + 16 * s + kh];
      const bf8 bv = *(const bf8 *)&Dt[(DT_OFF[L] + 32 * jb + col) * TP + 16 * s + kh];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    float *dst = p.part + (int64_t)blockIdx.x * PSTRIDE + bk * PBLK + lane;
    v16f old = zero16();
    if (!first)
#pragma unroll
      for (int r = 0; r < 16; r++) old[r] = dst[64 * r];
#pragma unroll
    for (int r = 0; r < 16; r++) __builtin_nontemporal_store(first ? acc[r] : old[r] + acc[r], &dst[64 * r]);
  }
}

// row `row`'s normalised inputs 3c8 .. 3c8+2 and target c8 (zero past n)
struct Inputs {
  float x[3], t;
};
__device__ __forceinline__ Inputs load_inputs(const RowsArgs &p, int row, int c8) {
  Inputs in = {{0.0f, 0.0f, 0.0f}, 0.0f};
  if (row < p.n) {
    if (p.idx || p.shuffled) {
      const int64_t s = p.shuffled ? shuffled_sample(p.seed, p.pos0 + row, p.nsamples) : p.idx[row];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const int k = 3 * c8 + c;
        in.x[c] = (p.xs[s * NIN + k] - p.norm[k]) / p.norm[NIN + k];
      }
      if (p.backward) in.t = (p.ys[s * NOUT + c8] - p.norm[2 * NIN + c8]) / p.norm[2 * NIN + NOUT + c8];
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) in.x[c] = p.x[(int64_t)row * NIN + 3 * c8 + c];
      if (p.backward) in.t = p.y[(int64_t)row * NOUT + c8];
    }
  }
  return in;
}

// the rows pass of workgroup blockIdx.x of gridDim.x (see the top of the file)
__device__ __forceinline__ void rows_pass(const RowsArgs &p, RowsLDS &sh) {
  __bf16(*act)[RW * AP] = sh.act;
  __bf16 *Ht = sh.Ht, *Dt = sh.Dt;
  float *bias = sh.bias, *tgt = sh.tgt;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, col = lane & 31;
  const int kh = 8 * (lane >> 5), kq = 4 * (lane >> 5);
  const int f = 32 * wave + col;       // the weight row this lane supplies
  // Loads are issued in the order they are needed: the biases and the first
  // block's inputs, then the weights layer by layer; the barriers wait for
  // LDS only, so each layer waits for its own weights and no store waits.
  float bv[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int e = tid + 64 * NWAVE * r, l = e < 3 * HID ? e / HID : 3;
    bv[r] = e < NBIAS ? p.w[(l == 0 ? LAYER_B[0] : l == 1 ? LAYER_B[1] : l == 2 ? LAYER_B[2] : LAYER_B[3]) + e -
                            l * HID]
                      : 0.0f;
  }
  const int nblk = (p.n + RW - 1) / RW;
  Inputs in = load_inputs(p, blockIdx.x * RW + (tid >> 3), tid & 7);
  // this wave's weight slices, for the whole launch
  bf8 w1[2], w2[8], w3[8], w4[8], v4[1], v3[8], v2[8];
#pragma unroll
  for (int s = 0; s < 2; s++) w1[s] = *(const bf8 *)&p.wT[WT_OFF[0] + f * K1 + 16 * s + kh];
#pragma unroll
  for (int s = 0; s < 8; s++) w2[s] = *(const bf8 *)&p.wT[WT_OFF[1] + f * HID + 16 * s + kh];
#pragma unroll
  for (int s = 0; s < 8; s++) w3[s] = *(const bf8 *)&p.wT[WT_OFF[2] + f * HID + 16 * s + kh];
  // the output layer is wave 0's alone (W4^T rows 8..31 are zero)
#pragma unroll
  for (int s = 0; s < 8; s++) w4[s] = *(const bf8 *)&p.wT[WT_OFF[3] + col * HID + 16 * s + kh];
  if (p.backward) {
    v4[0] = *(const bf8 *)&p.wB[WB_OFF[3] + f * N4B + kh];
#pragma unroll
    for (int s = 0; s < 8; s++) v3[s] = *(const bf8 *)&p.wB[WB_OFF[2] + f * HID + 16 * s + kh];
#pragma unroll
    for (int s = 0; s < 8; s++) v2[s] = *(const bf8 *)&p.wB[WB_OFF[1] + f * HID + 16 * s + kh];
  }
#pragma unroll
  for (int r = 0; r < 2; r++)
    if (tid + 64 * NWAVE * r < NBIAS) bias[tid + 64 * NWAVE * r] = bv[r];
  // the pads of layer 1's inputs and layer 4's gradients, never written below
  for (int e = tid; e < (K1 - NIN) * TP; e += 64 * NWAVE) Ht[(HT_OFF[0] + NIN) * TP + e] = (__bf16)0.0f;
  for (int e = tid; e < (32 - NOUT) * TP; e += 64 * NWAVE) Dt[(DT_OFF[3] + NOUT) * TP + e] = (__bf16)0.0f;
  const float sc = 2.0f / (float)(p.n * NOUT);
  float sq = 0.0f;                     // wave 0: squared errors of this workgroup's rows
  bool first = true;
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x, first = false) {
    const int r0 = blk * RW;
    if (!first) {
      // the last block's readers are done, and (a full barrier) its partial
      // gradients are stored before this block reads them back
      __syncthreads();
      in = load_inputs(p, r0 + (tid >> 3), tid & 7);
    }
    // ---- inputs: thread t takes row t/8 of the block, inputs 3(t%8) .. +2
    // and target t%8
    {
      const int rr = tid >> 3, c8 = tid & 7;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const __bf16 b = (__bf16)in.x[c];
        act[0][rr * AP + 3 * c8 + c] = b;
        Ht[(HT_OFF[0] + 3 * c8 + c) * TP + rr] = b;
      }
      act[0][rr * AP + NIN + c8] = (__bf16)0.0f;
      tgt[rr * NOUT + c8] = in.t;
    }
    lds_barrier();
    // ---- forward
    __bf16 *H1 = p.backward ? &Ht[HT_OFF[1] * TP] : nullptr;
    __bf16 *H2 = p.backward ? &Ht[HT_OFF[2] * TP] : nullptr;
    __bf16 *H3 = p.backward ? &Ht[HT_OFF[3] * TP] : nullptr;
    float sg1[16], sg2[16], sg3[16];
    forward_layer<2>(w1, act[0], act[1], &bias[0], sg1, H1, wave, col, kh, kq);
    lds_barrier();
    forward_layer<8>(w2, act[1], act[0], &bias[HID], sg2, H2, wave, col, kh, kq);
    lds_barrier();
    forward_layer<8>(w3, act[0], act[1], &bias[2 * HID], sg3, H3, wave, col, kh, kq);
    lds_barrier();
    if (wave == 0) {
      // the lane's register group 0 holds outputs kq .. kq+3 of row col
      const v16f acc = product<8>(w4, act[1], col, kh);
      const int row = r0 + col;
      const bool live = row < p.n;
      float y[4];
#pragma unroll
      for (int q = 0; q < 4; q++) y[q] = acc[q] + bias[3 * HID + kq + q];
      if (!p.backward) {
        if (live) *(float4 *)&p.out[(int64_t)row * NOUT + kq] = make_float4(y[0], y[1], y[2], y[3]);
      } else {
        // dY = 2 (y - t) / (n 8), the mean-squared error's gradient
        floats<4>::type dy;
        bf4 dv, zero;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const float d = live ? y[q] - tgt[col * NOUT + kq + q] : 0.0f;
          sq += d * d;
          dy[q] = sc * d;
          dv[q] = (__bf16)dy[q];
          zero[q] = (__bf16)0.0f;
          Dt[(DT_OFF[3] + kq + q) * TP + col] = dv[q];
        }
        *(bf4 *)&act[0][col * AP + kq] = dv;
        *(bf4 *)&act[0][col * AP + NOUT + kq] = zero;
        int vi;
        const float s = half_sums<4>(dy, lane, vi);
        if ((lane & 7) == 0) part_add(p, PB_OFF + 3 * HID + kq + vi, s, first);
      }
    }
    if (!p.backward) continue;
    lds_barrier();
    // ---- backward: layers 3, 2, 1, each beside the weight gradients of
    // the layer above, whose dZ the last barrier published (independent
    // products: the matrix cores run them while the epilogues run)
    weight_grad_blocks<3>(p, Ht, Dt, wave, first, lane, col, kh);
    backward_layer<1>(p, v4, act[0], act[1], sg3, &Dt[DT_OFF[2] * TP], PB_OFF + 2 * HID, first, wave, lane, col, kh, kq);
    lds_barrier();
    weight_grad_blocks<2>(p, Ht, Dt, wave, first, lane, col, kh);
    backward_layer<8>(p, v3, act[1], act[0], sg2, &Dt[DT_OFF[1] * TP], PB_OFF + HID, first, wave, lane, col, kh, kq);
    lds_barrier();
    weight_grad_blocks<1>(p, Ht, Dt, wave, first, lane, col, kh);
    backward_layer<8>(p, v2, act[0], nullptr, sg1, &Dt[DT_OFF[0] * TP], PB_OFF, first, wave, lane, col, kh,
                      kq);
    lds_barrier();
    weight_grad_blocks<0>(p, Ht, Dt, wave, first, lane, col, kh);
  }
  if (p.backward && wave == 0) {
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
    if (lane == 0) p.part[(int64_t)blockIdx.x * PSTRIDE + PLOSS] = sq;
  }
}

__global__ void __launch_bounds__(64 * NWAVE) __attribute__((amdgpu_waves_per_eu(1, 1)))
rows_kernel(RowsArgs p) {
  __shared__ RowsLDS sh;
  rows_pass(p, sh);
}

// (layer, input, output) of weight k of the packed vector; false for a bias
__device__ __forceinline__ bool weight_at(int k, int &l, int &i, int &j) {
#pragma unroll
  for (l = 0; l < 4; l++) {
    if (k < LAYER_B[l]) {
      const int e = k - LAYER_W[l];
      i = e / LAYER_OUT[l];
      j = e - i * LAYER_OUT[l];
      return true;
    }
    if (k < LAYER_B[l] + LAYER_OUT[l]) return false;
  }
  return false;
}

// where parameter k (k = NPARAM: the loss) sits in a slot of partials
__device__ __forceinline__ int partial_index(int k) {
#pragma unroll
  for (int l = 0; l < 4; l++) {
    const int N = LAYER_OUT[l];
    if (k < LAYER_B[l]) {
      const int e = k - LAYER_W[l], i = e / N, j = e - i * N;
      const int ib = i >> 5, jb = j >> 5, ii = i & 31;
      const int bk = l == 0 ? jb : l == 1 ? 4 + 4 * ib + jb : l == 2 ? 20 + 4 * ib + jb : 36 + ib;
      // D[8g + 4h + q][c] is register 4g + q of lane 32h + c
      return bk * PBLK + (4 * (ii >> 3) + (ii & 3)) * 64 + 32 * ((ii >> 2) & 1) + (j & 31);
    }
    if (k < LAYER_B[l] + N) return PB_OFF + l * HID + (k - LAYER_B[l]);
  }
  return PLOSS;
}

// the BF16 copies of weight (l, i, j)
__device__ __forceinline__ void store_copies(uint16_t *wT, uint16_t *wB, int l, int i, int j, float wij) {
  const uint16_t b = f32_to_bf16(wij);
  wT[WT_OFF[l] + j * (l == 0 ? K1 : HID) + i] = b;
  if (l > 0) wB[WB_OFF[l] + i * (l == 3 ? N4B : HID) + j] = b;
}

// parameter k's gradient from the partials, then its Adam step (or the
// gradient itself); k = NPARAM: the loss.  Four consecutive threads take
// one parameter: thread s sums partials s, s+4, ..., so that its loads are
// all in flight at once, and the four sums are added pairwise (the same
// bits in every thread of the four, whatever the order of arrival).
__global__ void __launch_bounds__(256) update_kernel(UpdateArgs u) {
  const int t = blockIdx.x * 256 + threadIdx.x, k = t >> 2, s4 = t & 3;
  if (k > NPARAM) return;              // (whole quads)
  const int loc = partial_index(k);
  float s = 0.0f;
#pragma unroll 16
  for (int w = s4; w < u.npart; w += 4) s += u.part[(int64_t)w * PSTRIDE + loc];
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  if (s4) return;
  if (k == NPARAM) {
    if (u.loss) *u.loss = s / (float)(u.n * NOUT);
    return;
  }
  if (!u.adam) {
    u.g[k] = s;
    return;
  }
  const float mi = u.b1 * u.m[k] + (1.0f - u.b1) * s;
  const float vi = u.b2 * u.v[k] + (1.0f - u.b2) * s * s;
  const float wk = u.w[k] - u.lr * (mi / u.c1) / (sqrtf(vi / u.c2) + u.eps);
  u.m[k] = mi;
  u.v[k] = vi;
  u.w[k] = wk;
  int l, i, j;
  if (weight_at(k, l, i, j)) store_copies(u.wT, u.wB, l, i, j, wk);
}

// the BF16 copies of every weight, from the FP32 parameters
__global__ void copies_kernel(const float *__restrict__ w, uint16_t *__restrict__ wT, uint16_t *__restrict__ wB) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  int l, i, j;
  if (k < NPARAM && weight_at(k, l, i, j)) store_copies(wT, wB, l, i, j, w[k]);
}

}  // namespace

// ---------------------------------------------------------------------------
// context

struct cable_mlp {
  int batch = 0;
  hipStream_t stream = nullptr;
  // parameters: FP32 master, gradient, Adam moments (NPARAM each), and the
  // BF16 copies in the two layouts the products read
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wT = nullptr, *wB = nullptr;
  float *norm = nullptr;  // mu_x[24], sd_x[24], mu_y[8], sd_y[8]
  // rows handed in by forward/gradient, the outputs, the loss
  float *x = nullptr, *y = nullptr, *out = nullptr, *loss = nullptr;
  float *part = nullptr;  // partial gradients, GMAX x PSTRIDE
  // training samples (device)
  float *xs = nullptr, *ys = nullptr;
  int64_t nsamples = 0;
  bool own_samples = false;
  int64_t *idx = nullptr;     // sample order of the batches being trained
  int64_t idxcap = 0;
  float *losses = nullptr;    // per-batch losses of the current train call
  int64_t losscap = 0;
  int64_t adam_t = 0;
};

namespace {

int groups(int n) { return std::min((n + RW - 1) / RW, GMAX); }

RowsArgs rows_args(cable_mlp *m, int n, bool backward, const int64_t *idx) {
  RowsArgs a;
  memset(&a, 0, sizeof a);
  a.xs = m->xs;
  a.ys = m->ys;
  a.idx = idx;
  a.x = m->x;
  a.y = m->y;
  a.norm = m->norm;
  a.w = m->w;
  a.wT = m->wT;
  a.wB = m->wB;
  a.out = m->out;
  a.part = m->part;
  a.n = n;
  a.backward = backward ? 1 : 0;
  return a;
}

// lr < 0: store the gradient; else Adam at rate lr (c1, c2 for the next step)
UpdateArgs update_args(cable_mlp *m, int n, float lr, float *loss) {
  UpdateArgs u;
  memset(&u, 0, sizeof u);
  u.part = m->part;
  u.npart = groups(n);
  u.n = n;
  u.w = m->w;
  u.m = m->m;
  u.v = m->v;
  u.g = m->g;
  u.wT = m->wT;
  u.wB = m->wB;
  u.loss = loss;
  if (lr >= 0.0f) {
    u.adam = 1;
    u.lr = lr;
    u.b1 = 0.9f;
    u.b2 = 0.999f;
    u.eps = 1e-8f;
    u.c1 = 1.0f - powf(u.b1, (float)(m->adam_t + 1));
    u.c2 = 1.0f - powf(u.b2, (float)(m->adam_t + 1));
  }
  return u;
}

// one pass of rows_kernel over n rows: forward only into m->out, or with
// the backward pass into the partials; idx: gather raw samples (else m->x, m->y)
int rows(cable_mlp *m, int n, bool backward, const int64_t *idx) {
  hipLaunchKernelGGL(rows_kernel, dim3(groups(n)), dim3(64 * NWAVE), 0, m->stream, rows_args(m, n, backward, idx));
  HIPCHECK(hipGetLastError());
  return CABLE_MLP_OK;
}

// after a backward pass over n rows: the gradient into m->g (lr < 0) or an
// Adam step at rate lr; the loss into *loss (device, optional)
int update(cable_mlp *m, int n, float lr, float *loss) {
  const UpdateArgs u = update_args(m, n, lr, loss);
  if (lr >= 0.0f) m->adam_t++;
  hipLaunchKernelGGL(update_kernel, dim3(((NPARAM + 1) * 4 + 255) / 256), dim3(256), 0, m->stream, u);
  HIPCHECK(hipGetLastError());
  return CABLE_MLP_OK;
}

// room for the losses of nbatches batches
int ensure_losses(cable_mlp *m, int64_t nbatches) {
  if (nbatches > m->losscap) {
    if (m->losses) HIPCHECK(hipFree(m->losses));
    m->losses = nullptr;
    m->losscap = 0;
    HIPCHECK(hipMalloc(&m->losses, sizeof(float) * (size_t)nbatches));
    m->losscap = nbatches;
  }
  return CABLE_MLP_OK;
}

int copy_rows_in(cable_mlp *m, const float *host, float *dev, int n, int cols) {
  HIPCHECK(hipMemcpyAsync(dev, host, sizeof(float) * (size_t)n * cols, hipMemcpyHostToDevice, m->stream));
  return CABLE_MLP_OK;
}

}  // namespace

extern "C" {

const char *cable_mlp_last_error(void) { return g_err.c_str(); }
int cable_mlp_nparam(void) { return NPARAM; }

int cable_mlp_create(cable_mlp **out, int batch) {
  if (!out || batch <= 0 || batch > CABLE_MLP_MAX_BATCH) return fail(CABLE_MLP_E_ARG, "cable_mlp_create: bad batch %d", batch);
  cable_mlp *m = new cable_mlp();
  m->batch = batch;
  hipError_t e = hipStreamCreate(&m->stream);
  auto al = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
    if (e == hipSuccess) e = hipMemset(*p, 0, bytes);
  };
  al((void **)&m->w, sizeof(float) * NPARAM);
  al((void **)&m->g, sizeof(float) * NPARAM);
  al((void **)&m->m, sizeof(float) * NPARAM);
  al((void **)&m->v, sizeof(float) * NPARAM);
  al((void **)&m->wT, sizeof(uint16_t) * WT_SIZE);   // the pads stay zero
  al((void **)&m->wB, sizeof(uint16_t) * WB_SIZE);
  al((void **)&m->norm, sizeof(float) * NNORM);
  al((void **)&m->x, sizeof(float) * (size_t)batch * NIN);
  al((void **)&m->y, sizeof(float) * (size_t)batch * NOUT);
  al((void **)&m->out, sizeof(float) * (size_t)batch * NOUT);
  al((void **)&m->loss, sizeof(float));
  al((void **)&m->part, sizeof(float) * (size_t)GMAX * PSTRIDE);
  if (e != hipSuccess) {
    cable_mlp_destroy(m);
    return fail(CABLE_MLP_E_HIP, "cable_mlp_create: %s", hipGetErrorString(e));
  }
  // identity normalisation until set
  std::vector<float> nrm(NNORM, 0.0f);
  for (int c = 0; c < NIN; c++) nrm[NIN + c] = 1.0f;
  for (int c = 0; c < NOUT; c++) nrm[2 * NIN + NOUT + c] = 1.0f;
  hipMemcpy(m->norm, nrm.data(), sizeof(float) * NNORM, hipMemcpyHostToDevice);
  *out = m;
  return CABLE_MLP_OK;
}

int cable_mlp_destroy(cable_mlp *m) {
  if (!m) return CABLE_MLP_OK;
  void *ps[] = {m->w, m->g, m->m, m->v, m->wT, m->wB, m->norm, m->x, m->y, m->out, m->loss, m->part, m->idx,
                m->losses};
  for (void *p : ps)
    if (p) hipFree(p);
  if (m->own_samples) {
    if (m->xs) hipFree(m->xs);
    if (m->ys) hipFree(m->ys);
  }
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return CABLE_MLP_OK;
}

int cable_mlp_set_params(cable_mlp *m, const float *params) {
  if (!m || !params) return fail(CABLE_MLP_E_ARG, "cable_mlp_set_params: null argument");
  HIPCHECK(hipMemcpyAsync(m->w, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(copies_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->w, m->wT, m->wB);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemsetAsync(m->m, 0, sizeof(float) * NPARAM, m->stream));
  HIPCHECK(hipMemsetAsync(m->v, 0, sizeof(float) * NPARAM, m->stream));
  m->adam_t = 0;
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_get_params(cable_mlp *m, float *params) {
  if (!m || !params) return fail(CABLE_MLP_E_ARG, "cable_mlp_get_params: null argument");
  HIPCHECK(hipMemcpyAsync(params, m->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_get_grads(cable_mlp *m, float *grads) {
  if (!m || !grads) return fail(CABLE_MLP_E_ARG, "cable_mlp_get_grads: null argument");
  HIPCHECK(hipMemcpyAsync(grads, m->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_set_norm(cable_mlp *m, const float *norm) {
  if (!m || !norm) return fail(CABLE_MLP_E_ARG, "cable_mlp_set_norm: null argument");
  for (int c = 0; c < NIN; c++)
    if (!(norm[NIN + c] > 0)) return fail(CABLE_MLP_E_ARG, "cable_mlp_set_norm: sd_x[%d] not positive", c);
  for (int c = 0; c < NOUT; c++)
    if (!(norm[2 * NIN + NOUT + c] > 0)) return fail(CABLE_MLP_E_ARG, "cable_mlp_set_norm: sd_y[%d] not positive", c);
  HIPCHECK(hipMemcpyAsync(m->norm, norm, sizeof(float) * NNORM, hipMemcpyHostToDevice, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_forward(cable_mlp *m, const float *x, int n, float *y) {
  if (!m || !x || !y || n <= 0 || n > m->batch) return fail(CABLE_MLP_E_ARG, "cable_mlp_forward: bad arguments");
  int r = copy_rows_in(m, x, m->x, n, NIN);
  if (!r) r = rows(m, n, false, nullptr);
  if (r) return r;
  HIPCHECK(hipMemcpyAsync(y, m->out, sizeof(float) * (size_t)n * NOUT, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_gradient(cable_mlp *m, const float *x, const float *t, int n, float *loss) {
  if (!m || !x || !t || n <= 0 || n > m->batch) return fail(CABLE_MLP_E_ARG, "cable_mlp_gradient: bad arguments");
  int r = copy_rows_in(m, x, m->x, n, NIN);
  if (!r) r = copy_rows_in(m, t, m->y, n, NOUT);
  if (!r) r = rows(m, n, true, nullptr);
  if (!r) r = update(m, n, -1.0f, m->loss);
  if (r) return r;
  if (loss) HIPCHECK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_load_samples(cable_mlp *m, const float *x, const float *y, int64_t n) {
  if (!m || !x || !y || n <= 0) return fail(CABLE_MLP_E_ARG, "cable_mlp_load_samples: bad arguments");
  if (m->own_samples) {
    hipFree(m->xs);
    hipFree(m->ys);
  }
  m->xs = m->ys = nullptr;
  HIPCHECK(hipMalloc(&m->xs, sizeof(float) * (size_t)n * NIN));
  HIPCHECK(hipMalloc(&m->ys, sizeof(float) * (size_t)n * NOUT));
  m->own_samples = true;
  HIPCHECK(hipMemcpyAsync(m->xs, x, sizeof(float) * (size_t)n * NIN, hipMemcpyHostToDevice, m->stream));
  HIPCHECK(hipMemcpyAsync(m->ys, y, sizeof(float) * (size_t)n * NOUT, hipMemcpyHostToDevice, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  m->nsamples = n;
  return CABLE_MLP_OK;
}

int cable_mlp_use_samples(cable_mlp *m, const float *xs_dev, const float *ys_dev, int64_t n) {
  if (!m || !xs_dev || !ys_dev || n <= 0) return fail(CABLE_MLP_E_ARG, "cable_mlp_use_samples: bad arguments");
  if (m->own_samples) {
    hipFree(m->xs);
    hipFree(m->ys);
    m->own_samples = false;
  }
  m->xs = (float *)xs_dev;
  m->ys = (float *)ys_dev;
  m->nsamples = n;
  return CABLE_MLP_OK;
}

int cable_mlp_shuffled_order(uint64_t seed, int64_t n, int64_t first, int64_t count, int64_t *out) {
  if (n <= 0 || first < 0 || count < 0 || (count && !out))
    return fail(CABLE_MLP_E_ARG, "cable_mlp_shuffled_order: bad arguments");
  for (int64_t i = 0; i < count; i++) out[i] = shuffled_sample(seed, first + i, n);
  return CABLE_MLP_OK;
}

int cable_mlp_train_shuffled(cable_mlp *m, uint64_t seed, int64_t first_batch, int64_t nbatches, float lr,
                             float *losses) {
  if (!m || first_batch < 0 || nbatches <= 0) return fail(CABLE_MLP_E_ARG, "cable_mlp_train_shuffled: bad arguments");
  if (!m->xs) return fail(CABLE_MLP_E_STATE, "cable_mlp_train_shuffled: no samples loaded");
  int r = ensure_losses(m, nbatches);
  if (r) return r;
  const int B = m->batch;
  for (int64_t b = 0; b < nbatches; b++) {
    RowsArgs a = rows_args(m, B, true, nullptr);
    a.shuffled = 1;
    a.seed = seed;
    a.nsamples = m->nsamples;
    a.pos0 = (first_batch + b) * B;
    hipLaunchKernelGGL(rows_kernel, dim3(groups(B)), dim3(64 * NWAVE), 0, m->stream, a);
    HIPCHECK(hipGetLastError());
    r = update(m, B, lr, m->losses + b);
    if (r) return r;
  }
  if (losses) HIPCHECK(hipMemcpyAsync(losses, m->losses, sizeof(float) * (size_t)nbatches, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_train(cable_mlp *m, const int64_t *order, int64_t nbatches, float lr, float *losses) {
  if (!m || !order || nbatches <= 0) return fail(CABLE_MLP_E_ARG, "cable_mlp_train: bad arguments");
  if (!m->xs) return fail(CABLE_MLP_E_STATE, "cable_mlp_train: no samples loaded");
  const int B = m->batch;
  const int64_t total = nbatches * B;
  for (int64_t i = 0; i < total; i++)
    if (order[i] < 0 || order[i] >= m->nsamples)
      return fail(CABLE_MLP_E_ARG, "cable_mlp_train: sample index %lld out of range", (long long)order[i]);
  if (total > m->idxcap) {
    if (m->idx) HIPCHECK(hipFree(m->idx));
    m->idx = nullptr;
    m->idxcap = 0;
    HIPCHECK(hipMalloc(&m->idx, sizeof(int64_t) * (size_t)total));
    m->idxcap = total;
  }
  int r0 = ensure_losses(m, nbatches);
  if (r0) return r0;
  HIPCHECK(hipMemcpyAsync(m->idx, order, sizeof(int64_t) * (size_t)total, hipMemcpyHostToDevice, m->stream));
  // two launches per batch.  (One cooperative launch for all batches, with
  // grid barriers between the passes, measured three times slower.)
  for (int64_t b = 0; b < nbatches; b++) {
    int r = rows(m, B, true, m->idx + b * B);
    if (!r) r = update(m, B, lr, m->losses + b);
    if (r) return r;
  }
  if (losses) HIPCHECK(hipMemcpyAsync(losses, m->losses, sizeof(float) * (size_t)nbatches, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

}  // extern "C"
