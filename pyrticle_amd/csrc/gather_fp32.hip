// gather_fp32.hip -- instantiations of the FP32-compensated deposition gather
// on the matrix cores (k_gather_mf, two waves per target element: the fastest
// of 1-8 measured on the cylinder workload).
#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP, int W>
static void launch_mf(const GatherArgs &ga, int K, bool a2) {
  dim3 grid(gather_grid(K)), block(64 * W);
  hipStream_t s = stream();
  if (a2) hipLaunchKernelGGL((k_gather_mf<D, NP, W, true>), grid, block, 0, s, ga);
  else hipLaunchKernelGGL((k_gather_mf<D, NP, W, false>), grid, block, 0, s, ga);
  PYR_CHECK_LAUNCH();
}

void launch_gather_fp32(int D, int Np, const GatherArgs &ga, int K, bool a2, bool count) {
  // the matrix-core form evaluates every candidate pair; it has no pair
  // counters (count mode is served by the FP64 gather)
  if (count) throw Error(PYR_ERR_INVALID_ARG, "deposit counting needs the FP64 precision");
  if (D == 2) {
    switch (Np) {
      case 3: return launch_mf<2, 3, 2>(ga, K, a2);
      case 6: return launch_mf<2, 6, 2>(ga, K, a2);
      case 10: return launch_mf<2, 10, 2>(ga, K, a2);
      case 15: return launch_mf<2, 15, 2>(ga, K, a2);
    }
  } else {
    switch (Np) {
      case 4: return launch_mf<3, 4, 2>(ga, K, a2);
      case 10: return launch_mf<3, 10, 2>(ga, K, a2);
      case 20: return launch_mf<3, 20, 2>(ga, K, a2);
      case 35: return launch_mf<3, 35, 2>(ga, K, a2);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

}  // namespace pyr
