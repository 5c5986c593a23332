// cbm_kernels_mixed.hip -- kernels of the mixed path: radiation, roughness and
// the canopy iteration in FP32; the soil water and heat state, the two
// tridiagonal solves and the balance accumulators in FP64.  Compiled with FMA contraction inside expressions and the
// hardware approximations of the elementary functions: the path is held to
// the flux tolerance of its contract, not to the reference's rounding.
#include "cbm_kernel.h"

namespace cbm {

hipError_t launch_mixed(bool full, const StepArgs &a, int64_t mp, hipStream_t stream) {
  if (full) return launch_cbm<float, true, false>(a, mp, stream);
  return a.cons_part ? launch_cbm<float, false, true, MODE_PLAIN, true>(a, mp, stream)
                     : launch_cbm<float, false, true>(a, mp, stream);
}

hipError_t launch_mixed_mode(int mode, const StepArgs &a, int64_t mp, hipStream_t stream) {
  switch (mode) {
    case MODE_CAPTURE: return launch_cbm<float, false, true, MODE_CAPTURE>(a, mp, stream);
    case MODE_HYBRID: return launch_cbm<float, false, true, MODE_HYBRID>(a, mp, stream);
    case MODE_UNLIT: return launch_cbm<float, false, true, MODE_UNLIT>(a, mp, stream);
    default: return launch_cbm<float, false, true, MODE_PLAIN>(a, mp, stream);
  }
}

hipError_t launch_mlp_infer(const StepArgs &a, const uint16_t *wt, const float *params, float *pred, float *xraw,
                            int64_t mp, hipStream_t stream) {
  // one workgroup per CU (the kernel's LDS holds one), each looping over
  // every gridDim.x-th block of 256 tiles
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  const int64_t blocks = (mp + infer::TILES - 1) / infer::TILES;
  const int64_t grid = blocks < cus ? blocks : cus;
  hipLaunchKernelGGL(mlp_infer_kernel<infer::WAVES>, dim3((unsigned)grid), dim3(64 * infer::WAVES), 0, stream, a, wt,
                     params, pred, xraw);
  return hipGetLastError();
}

}  // namespace cbm
