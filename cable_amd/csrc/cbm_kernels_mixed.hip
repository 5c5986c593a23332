// cbm_kernels_mixed.hip -- kernels of the mixed path: default-REAL quantities in FP32, REAL(r_2) and the
// solves in FP64.  Compiled with FMA contraction inside expressions and the
// hardware approximations of the elementary functions: the path is held to
// the flux tolerance of its contract, not to the reference's rounding.
#include "cbm_kernel.h"

namespace cbm {

hipError_t launch_mixed(bool full, const StepArgs &a, int64_t mp, hipStream_t stream) {
  return full ? launch_cbm<float, true, false>(a, mp, stream) : launch_cbm<float, false, true>(a, mp, stream);
}

}  // namespace cbm
