// average.hip -- instantiations of the shape-weighted field average of the
// averaging pusher (k_shape_average), a per-particle gather over the
// transposed candidate lists.
#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP>
static void launch_average(const AverageArgs &ga, unsigned grid, bool a2, bool vis) {
  dim3 block(64);
  hipStream_t s = stream();
  if (vis) {
    if (a2) hipLaunchKernelGGL((k_shape_average<D, NP, true, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_shape_average<D, NP, false, true>), grid, block, 0, s, ga);
  } else {
    if (a2) hipLaunchKernelGGL((k_shape_average<D, NP, true, false>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_shape_average<D, NP, false, false>), grid, block, 0, s, ga);
  }
  PYR_CHECK_LAUNCH();
}

void launch_shape_average(int D, int Np, const AverageArgs &ga, unsigned grid, bool a2,
                          bool vis) {
  if (D == 2) {
    switch (Np) {
      case 3: return launch_average<2, 3>(ga, grid, a2, vis);
      case 6: return launch_average<2, 6>(ga, grid, a2, vis);
      case 10: return launch_average<2, 10>(ga, grid, a2, vis);
      case 15: return launch_average<2, 15>(ga, grid, a2, vis);
    }
  } else {
    switch (Np) {
      case 4: return launch_average<3, 4>(ga, grid, a2, vis);
      case 10: return launch_average<3, 10>(ga, grid, a2, vis);
      case 20: return launch_average<3, 20>(ga, grid, a2, vis);
      case 35: return launch_average<3, 35>(ga, grid, a2, vis);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "the averaging pusher supports P1..P4 elements");
}

}  // namespace pyr
