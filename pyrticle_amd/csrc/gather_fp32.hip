// gather_fp32.hip -- instantiations of the FP32-compensated deposition gather:
// the matrix-core form k_gather_mf (default; two waves per target element,
// the fastest of 1-8 measured on the cylinder workload) and the VALU form
// k_gather<.., FP32>
// (PYR_GATHER_MFMA=0, kept for comparison; it reads the FP64 staging).
#include <cstdlib>

#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP, int NPW>
static void launch_valu(const GatherArgs &ga, int K, bool a2, bool count) {
  constexpr int G = (NP + NPW - 1) / NPW;
  dim3 grid(gather_grid(K)), block(64 * G);
  hipStream_t s = stream();
  if (a2) {
    if (count) hipLaunchKernelGGL((k_gather<D, NP, NPW, true, true, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_gather<D, NP, NPW, true, true, false>), grid, block, 0, s, ga);
  } else {
    if (count) hipLaunchKernelGGL((k_gather<D, NP, NPW, true, false, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_gather<D, NP, NPW, true, false, false>), grid, block, 0, s, ga);
  }
  PYR_CHECK_LAUNCH();
}

template <int D, int NP, int W>
static void launch_mf(const GatherArgs &ga, int K, bool a2) {
  dim3 grid(gather_grid(K)), block(64 * W);
  hipStream_t s = stream();
  if (a2) hipLaunchKernelGGL((k_gather_mf<D, NP, W, true>), grid, block, 0, s, ga);
  else hipLaunchKernelGGL((k_gather_mf<D, NP, W, false>), grid, block, 0, s, ga);
  PYR_CHECK_LAUNCH();
}

static int env_int(const char *name, int dflt) {
  const char *env = std::getenv(name);
  return env ? std::atoi(env) : dflt;
}

bool gather_fp32_uses_mfma() {
  static const bool on = env_int("PYR_GATHER_MFMA", 1) != 0;
  return on;
}

void launch_gather_fp32(int D, int Np, const GatherArgs &ga, int K, bool a2, bool count) {
  if (!gather_fp32_uses_mfma()) {
    // the VALU form's squares take t <= 1 (shape_unit_square)
    const bool unit = a2 && ga.R <= 0.5;
    if (D == 3 && Np == 20) return launch_valu<3, 20, 5>(ga, K, unit, count);
    if (D == 2 && Np == 10) return launch_valu<2, 10, 10>(ga, K, unit, count);
    throw Error(PYR_ERR_INVALID_ARG, "the VALU FP32 gather is built for P3 elements only");
  }
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
      case 20:
        return launch_mf<3, 20, 2>(ga, K, a2);
      case 35: return launch_mf<3, 35, 2>(ga, K, a2);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

}  // namespace pyr
