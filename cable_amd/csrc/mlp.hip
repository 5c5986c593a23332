// mlp.hip -- the canopy surrogate: a 24 -> 128 -> 128 -> 128 -> 8 MLP with
// SiLU, BF16 operands on the matrix cores (MFMA), FP32 accumulation, FP32
// master weights and Adam.  All kernels are hand-written for gfx950 wave64.
//
// Matrix products use v_mfma_f32_32x32x8_bf16: one wave computes a 32x32
// block of C = A.B, stepping K by 8.  Lane l holds
//   A[i][k] for i = l % 32, k = 4*(l/32) + 0..3
//   B[k][j] for j = l % 32, k = 4*(l/32) + 0..3
//   C[i][j] for j = l % 32, i = (r % 4) + 4*(l/32) + 8*(r/4), r = 0..15.
// Operands are read through strides, so the same kernel forms X.W, dZ.W^T
// and H^T.dZ; every element outside [0,M) x [0,K) x [0,N) reads as zero.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

enum Epi { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_BIAS = 2, EPI_SILU_GRAD = 3 };

// C[M x N] = A[M x K] . B[K x N]; one wave per 32x32 block of C.  With
// kchunk < K the product is split along K: block z sums k in
// [z*kchunk, (z+1)*kchunk) and stores its partial C at C + z*M*N (EPI_STORE
// only); reduce_kernel adds the partials in order.
template <class TA, class TB>
__global__ void __launch_bounds__(64) gemm_kernel(const TA *__restrict__ A, int64_t sam, int64_t sak,
                                                  const TB *__restrict__ B, int64_t sbk, int64_t sbn,
                                                  float *__restrict__ C, float *__restrict__ H, int M,
                                                  int N, int K, const float *__restrict__ bias,
                                                  const float *__restrict__ Zsaved, int epi, int kchunk) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int ia = i0 + (lane & 31), jb = j0 + (lane & 31), kq = 4 * (lane >> 5);
  const int kbeg = blockIdx.z * kchunk, kend = kbeg + kchunk < K ? kbeg + kchunk : K;
  C += (int64_t)blockIdx.z * M * N;
  v16f acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += 8) {
    v4s a, b;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int k = k0 + kq + e;
      const float av = (ia < M && k < kend) ? to_f32(A[(int64_t)ia * sam + (int64_t)k * sak]) : 0.0f;
      const float bv = (jb < N && k < kend) ? to_f32(B[(int64_t)k * sbk + (int64_t)jb * sbn]) : 0.0f;
      a[e] = (short)f32_to_bf16(av);
      b[e] = (short)f32_to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int j = j0 + (lane & 31);
  if (j >= N) return;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = i0 + (r & 3) + 4 * (lane >> 5) + 8 * (r >> 2);
    if (i >= M) continue;
    const int64_t o = (int64_t)i * N + j;
    float v = acc[r];
    switch (epi) {
      case EPI_BIAS_SILU: {
        v += bias[j];
        C[o] = v;             // pre-activation, kept for the backward pass
        H[o] = silu(v);
        break;
      }
      case EPI_BIAS: C[o] = v + bias[j]; break;
      case EPI_SILU_GRAD: C[o] = v * silu_grad(Zsaved[o]); break;
      default: C[o] = v; break;
    }
  }
}

// out[i] = sum_z part[z][i] for i < n, z < nz (split-K partials, in order)
__global__ void reduce_kernel(const float *__restrict__ part, int nz, int64_t n, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int z = 0; z < nz; z++) s += part[z * n + i];
  out[i] = s;
}

// column sums of dZ in row blocks: block (x, y) sums rows [y*rows, (y+1)*rows)
// of 32 columns into part[y][j] (eight row groups combined through LDS);
// reduce_kernel then adds the row blocks in order
__global__ void __launch_bounds__(256) colsum_kernel(const float *__restrict__ dZ, int M, int N, int rows,
                                                     float *__restrict__ part) {
  __shared__ float acc[8][32];
  const int c = threadIdx.x & 31, r = threadIdx.x >> 5, j = blockIdx.x * 32 + c;
  const int i0 = blockIdx.y * rows, i1 = i0 + rows < M ? i0 + rows : M;
  float s = 0.0f;
  if (j < N)
    for (int i = i0 + r; i < i1; i += 8) s += dZ[(int64_t)i * N + j];
  acc[r][c] = s;
  __syncthreads();
  if (r == 0 && j < N) {
    float t = 0.0f;
    for (int q = 0; q < 8; q++) t += acc[q][c];
    part[(int64_t)blockIdx.y * N + j] = t;
  }
}

// normalised batch: x = (raw[idx] - mu) / sd, y likewise; rows past n are 0
__global__ void gather_kernel(const float *__restrict__ xs, const float *__restrict__ ys,
                              const int64_t *__restrict__ idx, int n, float *__restrict__ xb,
                              float *__restrict__ yb, const float *__restrict__ norm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = idx[i];
#pragma unroll
  for (int c = 0; c < NIN; c++) xb[(int64_t)i * NIN + c] = (xs[s * NIN + c] - norm[c]) / norm[NIN + c];
#pragma unroll
  for (int c = 0; c < NOUT; c++)
    yb[(int64_t)i * NOUT + c] = (ys[s * NOUT + c] - norm[2 * NIN + c]) / norm[2 * NIN + NOUT + c];
}

// mean-squared-error gradient dY = 2 (Y - T) / (n * NOUT), and the loss
__global__ void mse_kernel(const float *__restrict__ Y, const float *__restrict__ T, int n,
                           float *__restrict__ dY, float *__restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n * NOUT) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d / (float)(n * NOUT);
    l = d * d / (float)(n * NOUT);
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// Adam on the FP32 master parameters; writes the BF16 working copy
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, uint16_t *__restrict__ wb, int n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = f32_to_bf16(wi);
}

__global__ void to_bf16_kernel(const float *__restrict__ w, uint16_t *__restrict__ wb, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f32_to_bf16(w[i]);
}

}  // namespace

// ---------------------------------------------------------------------------
// context

struct cable_mlp {
  int batch = 0;
  hipStream_t stream = nullptr;
  // parameters: FP32 master, BF16 copy, gradient, Adam moments (NPARAM each)
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wb = nullptr;
  float *norm = nullptr;  // mu_x[24], sd_x[24], mu_y[8], sd_y[8]
  // activations of one batch
  float *x = nullptr, *y = nullptr, *z[3] = {nullptr, nullptr, nullptr}, *h[3] = {nullptr, nullptr, nullptr};
  float *out = nullptr, *dy = nullptr, *dh = nullptr, *dz = nullptr, *loss = nullptr;
  float *part = nullptr;      // split-K partials of the weight gradients
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

template <class TA, class TB>
hipError_t gemm(hipStream_t s, const TA *A, int64_t sam, int64_t sak, const TB *B, int64_t sbk, int64_t sbn,
                float *C, float *H, int M, int N, int K, const float *bias, const float *Z, int epi,
                int nsplit = 1) {
  const int kchunk = ((K + nsplit - 1) / nsplit + 7) / 8 * 8;
  dim3 grid((M + 31) / 32, (N + 31) / 32, nsplit);
  hipLaunchKernelGGL((gemm_kernel<TA, TB>), grid, dim3(64), 0, s, A, sam, sak, B, sbk, sbn, C, H, M, N, K, bias,
                     Z, epi, kchunk);
  return hipGetLastError();
}

// weight gradients: K (the batch) split in chunks of KSPLIT_ROWS rows
constexpr int KSPLIT_ROWS = 256;

// forward pass of n rows of m->x into m->out (pre-activations kept)
int forward(cable_mlp *m, int n) {
  const uint16_t *W[4];
  const float *b[4];
  for (int l = 0; l < 4; l++) {
    W[l] = m->wb + LAYER_W[l];
    b[l] = m->w + LAYER_B[l];
  }
  const float *in = m->x;
  for (int l = 0; l < 3; l++) {
    const int K = LAYER_IN[l], N = LAYER_OUT[l];
    HIPCHECK(gemm(m->stream, in, (int64_t)K, (int64_t)1, W[l], (int64_t)N, (int64_t)1, m->z[l], m->h[l], n, N, K,
                  b[l], (const float *)nullptr, EPI_BIAS_SILU));
    in = m->h[l];
  }
  HIPCHECK(gemm(m->stream, in, (int64_t)HID, (int64_t)1, W[3], (int64_t)NOUT, (int64_t)1, m->out, (float *)nullptr,
                n, NOUT, HID, b[3], (const float *)nullptr, EPI_BIAS));
  return CABLE_MLP_OK;
}

// gradients of the batch loss w.r.t. every parameter, into m->g
int backward(cable_mlp *m, int n, float *loss) {
  HIPCHECK(hipMemsetAsync(loss, 0, sizeof(float), m->stream));
  hipLaunchKernelGGL(mse_kernel, dim3((n * NOUT + 255) / 256), dim3(256), 0, m->stream, m->out, m->y, n, m->dy,
                     loss);
  HIPCHECK(hipGetLastError());
  // layer 4: dW4 = H3^T dY, db4 = colsum dY, dH3 = dY W4^T
  const float *dcur = m->dy;
  int Ncur = NOUT;
  for (int l = 3; l >= 0; l--) {
    const int K = LAYER_IN[l], N = LAYER_OUT[l];
    const float *hin = l == 0 ? m->x : m->h[l - 1];
    // dW[l] (K x N) = hin^T (K x n) . dcur (n x N), split along the batch
    const int nsplit = (n + KSPLIT_ROWS - 1) / KSPLIT_ROWS;
    HIPCHECK(gemm(m->stream, hin, (int64_t)1, (int64_t)K, dcur, (int64_t)N, (int64_t)1, m->part,
                  (float *)nullptr, K, N, n, (const float *)nullptr, (const float *)nullptr, EPI_STORE, nsplit));
    hipLaunchKernelGGL(reduce_kernel, dim3((K * N + 255) / 256), dim3(256), 0, m->stream, m->part, nsplit,
                       (int64_t)K * N, m->g + LAYER_W[l]);
    HIPCHECK(hipGetLastError());
    // db[l] = column sums of dcur; the partials reuse m->part (nsplit x N <= its size)
    hipLaunchKernelGGL(colsum_kernel, dim3((N + 31) / 32, nsplit), dim3(256), 0, m->stream, dcur, n, N,
                       KSPLIT_ROWS, m->part);
    hipLaunchKernelGGL(reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, m->stream, m->part, nsplit, (int64_t)N,
                       m->g + LAYER_B[l]);
    HIPCHECK(hipGetLastError());
    if (l == 0) break;
    // dZ[l-1] = (dcur . W[l]^T) * silu'(Z[l-1]); W[l] is K x N row-major
    float *dzn = (dcur == m->dz) ? m->dh : m->dz;
    HIPCHECK(gemm(m->stream, dcur, (int64_t)N, (int64_t)1, m->wb + LAYER_W[l], (int64_t)1, (int64_t)N, dzn,
                  (float *)nullptr, n, K, N, (const float *)nullptr, m->z[l - 1], EPI_SILU_GRAD));
    dcur = dzn;
    Ncur = K;
  }
  (void)Ncur;
  return CABLE_MLP_OK;
}

int adam(cable_mlp *m, float lr) {
  m->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)m->adam_t), c2 = 1.0f - powf(b2, (float)m->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->w, m->m, m->v, m->g,
                     m->wb, NPARAM, lr, b1, b2, eps, c1, c2);
  HIPCHECK(hipGetLastError());
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
  al((void **)&m->wb, sizeof(uint16_t) * NPARAM);
  al((void **)&m->norm, sizeof(float) * NNORM);
  al((void **)&m->x, sizeof(float) * (size_t)batch * NIN);
  al((void **)&m->y, sizeof(float) * (size_t)batch * NOUT);
  for (int l = 0; l < 3; l++) {
    al((void **)&m->z[l], sizeof(float) * (size_t)batch * HID);
    al((void **)&m->h[l], sizeof(float) * (size_t)batch * HID);
  }
  al((void **)&m->out, sizeof(float) * (size_t)batch * NOUT);
  al((void **)&m->dy, sizeof(float) * (size_t)batch * NOUT);
  al((void **)&m->dh, sizeof(float) * (size_t)batch * HID);
  al((void **)&m->dz, sizeof(float) * (size_t)batch * HID);
  al((void **)&m->loss, sizeof(float));
  al((void **)&m->part, sizeof(float) * HID * HID * (size_t)((batch + KSPLIT_ROWS - 1) / KSPLIT_ROWS));
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
  void *ps[] = {m->w, m->g, m->m, m->v, m->wb, m->norm, m->x, m->y, m->z[0], m->z[1], m->z[2], m->h[0], m->h[1],
                m->h[2], m->out, m->dy, m->dh, m->dz, m->loss, m->idx, m->losses, m->part};
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
  hipLaunchKernelGGL(to_bf16_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->w, m->wb, NPARAM);
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
  if (!r) r = forward(m, n);
  if (r) return r;
  HIPCHECK(hipMemcpyAsync(y, m->out, sizeof(float) * (size_t)n * NOUT, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

int cable_mlp_gradient(cable_mlp *m, const float *x, const float *t, int n, float *loss) {
  if (!m || !x || !t || n <= 0 || n > m->batch) return fail(CABLE_MLP_E_ARG, "cable_mlp_gradient: bad arguments");
  int r = copy_rows_in(m, x, m->x, n, NIN);
  if (!r) r = copy_rows_in(m, t, m->y, n, NOUT);
  if (!r) r = forward(m, n);
  if (!r) r = backward(m, n, m->loss);
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
  if (nbatches > m->losscap) {
    if (m->losses) HIPCHECK(hipFree(m->losses));
    m->losses = nullptr;
    m->losscap = 0;
    HIPCHECK(hipMalloc(&m->losses, sizeof(float) * (size_t)nbatches));
    m->losscap = nbatches;
  }
  HIPCHECK(hipMemcpyAsync(m->idx, order, sizeof(int64_t) * (size_t)total, hipMemcpyHostToDevice, m->stream));
  for (int64_t b = 0; b < nbatches; b++) {
    hipLaunchKernelGGL(gather_kernel, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->xs, m->ys, m->idx + b * B, B,
                       m->x, m->y, m->norm);
    HIPCHECK(hipGetLastError());
    int r = forward(m, B);
    if (!r) r = backward(m, B, m->losses + b);
    if (!r) r = adam(m, lr);
    if (r) return r;
  }
  if (losses) HIPCHECK(hipMemcpyAsync(losses, m->losses, sizeof(float) * (size_t)nbatches, hipMemcpyDeviceToHost, m->stream));
  HIPCHECK(hipStreamSynchronize(m->stream));
  return CABLE_MLP_OK;
}

}  // extern "C"
