// cbm_kernels_fp64.hip -- kernels of the plain path: FP64 throughout, no FMA contraction, IEEE division
// and square root, so that each step reproduces the reference's rounding
// up to the device math library.
#include "cbm_kernel.h"

namespace cbm {

hipError_t launch_fp64(bool full, const StepArgs &a, int64_t mp, hipStream_t stream) {
  if (full) return launch_cbm<double, true, false>(a, mp, stream);
  return a.cons_part ? launch_cbm<double, false, true, MODE_PLAIN, true>(a, mp, stream)
                     : launch_cbm<double, false, true>(a, mp, stream);
}

}  // namespace cbm
