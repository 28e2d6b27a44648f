// deposit_adj.hip -- instantiations of the deposition adjoint (k_deposit_adj)
// and of the per-particle shape integral of the normalized deposition
// (k_shape_integral), both per-particle gathers over the transposed candidate
// lists.
#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP>
static void launch(const AdjArgs &ga, unsigned grid, int precision, bool a2) {
  dim3 block(64);
  hipStream_t s = stream();
  if (precision == PYR_FP64) {
    if (a2) hipLaunchKernelGGL((k_deposit_adj<D, NP, false, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_deposit_adj<D, NP, false, false>), grid, block, 0, s, ga);
  } else {
    if (a2) hipLaunchKernelGGL((k_deposit_adj<D, NP, true, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_deposit_adj<D, NP, true, false>), grid, block, 0, s, ga);
  }
  PYR_CHECK_LAUNCH();
}

void launch_deposit_adj(int D, int Np, const AdjArgs &ga, unsigned grid, int precision,
                        bool a2) {
  if (D == 2) {
    switch (Np) {
      case 3: return launch<2, 3>(ga, grid, precision, a2);
      case 6: return launch<2, 6>(ga, grid, precision, a2);
      case 10: return launch<2, 10>(ga, grid, precision, a2);
      case 15: return launch<2, 15>(ga, grid, precision, a2);
    }
  } else {
    switch (Np) {
      case 4: return launch<3, 4>(ga, grid, precision, a2);
      case 10: return launch<3, 10>(ga, grid, precision, a2);
      case 20: return launch<3, 20>(ga, grid, precision, a2);
      case 35: return launch<3, 35>(ga, grid, precision, a2);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

template <int D, int NP>
static void launch_integral(const IntegralArgs &ga, unsigned grid, bool a2) {
  dim3 block(64);
  if (a2) hipLaunchKernelGGL((k_shape_integral<D, NP, true>), grid, block, 0, stream(), ga);
  else hipLaunchKernelGGL((k_shape_integral<D, NP, false>), grid, block, 0, stream(), ga);
  PYR_CHECK_LAUNCH();
}

void launch_shape_integral(int D, int Np, const IntegralArgs &ga, unsigned grid, bool a2) {
  if (D == 2) {
    switch (Np) {
      case 3: return launch_integral<2, 3>(ga, grid, a2);
      case 6: return launch_integral<2, 6>(ga, grid, a2);
      case 10: return launch_integral<2, 10>(ga, grid, a2);
      case 15: return launch_integral<2, 15>(ga, grid, a2);
    }
  } else {
    switch (Np) {
      case 4: return launch_integral<3, 4>(ga, grid, a2);
      case 10: return launch_integral<3, 10>(ga, grid, a2);
      case 20: return launch_integral<3, 20>(ga, grid, a2);
      case 35: return launch_integral<3, 35>(ga, grid, a2);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

}  // namespace pyr
