// gather_fp64.hip -- instantiations of the FP64 deposition gather (k_gather).
#include <cstdlib>

#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP, int NPW>
static void launch(const GatherArgs &ga, int K, bool a2, bool count) {
  constexpr int G = (NP + NPW - 1) / NPW;
  dim3 grid(gather_grid(K)), block(64 * G);
  hipStream_t s = stream();
  if (a2) {
    if (count) hipLaunchKernelGGL((k_gather<D, NP, NPW, false, true, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_gather<D, NP, NPW, false, true, false>), grid, block, 0, s, ga);
  } else {
    if (count) hipLaunchKernelGGL((k_gather<D, NP, NPW, false, false, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_gather<D, NP, NPW, false, false, false>), grid, block, 0, s, ga);
  }
  PYR_CHECK_LAUNCH();
}

// Nodes per wave of the 3D gather (launch-shape tuning; the result does not
// depend on it beyond summation order across waves, which is fixed).
static int gather_npw(int dflt) {
  const char *env = std::getenv("PYR_GATHER_NPW");
  return env ? std::atoi(env) : dflt;
}

void launch_gather_fp64(int D, int Np, const GatherArgs &ga, int K, bool a2, bool count) {
  if (D == 2) {
    switch (Np) {
      case 3: return launch<2, 3, 3>(ga, K, a2, count);
      case 6: return launch<2, 6, 6>(ga, K, a2, count);
      case 10: return launch<2, 10, 10>(ga, K, a2, count);
      case 15: return launch<2, 15, 15>(ga, K, a2, count);
    }
  } else {
    switch (Np) {
      case 4: return launch<3, 4, 4>(ga, K, a2, count);
      case 10: return launch<3, 10, 10>(ga, K, a2, count);
      case 20:
        switch (gather_npw(10)) {
          case 5: return launch<3, 20, 5>(ga, K, a2, count);
          case 20: return launch<3, 20, 20>(ga, K, a2, count);
          default: return launch<3, 20, 10>(ga, K, a2, count);
        }
      case 35:
        switch (gather_npw(12)) {
          case 9: return launch<3, 35, 9>(ga, K, a2, count);
          default: return launch<3, 35, 12>(ga, K, a2, count);
        }
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

}  // namespace pyr
