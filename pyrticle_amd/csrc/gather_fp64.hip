// gather_fp64.hip -- instantiations of the FP64 deposition gathers: the
// list gather on the FP64 matrix cores (k_gather_mf64) and the candidate
// gather of the pair-counting mode (k_gather; nodes per wave: the whole
// element in 2D, 10 (P3) and 12 (P4) in 3D, the fastest measured on the
// cylinder workloads; the result does not depend on it beyond the fixed
// summation order across waves).
#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP, int NPW>
static void launch(const GatherArgs &ga, int K, bool a2, bool count) {
  constexpr int G = (NP + NPW - 1) / NPW;
  dim3 grid(gather_grid(K)), block(64 * G);
  hipStream_t s = stream();
  if (a2) {
    if (count) hipLaunchKernelGGL((k_gather<D, NP, NPW, true, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_gather<D, NP, NPW, true, false>), grid, block, 0, s, ga);
  } else {
    if (count) hipLaunchKernelGGL((k_gather<D, NP, NPW, false, true>), grid, block, 0, s, ga);
    else hipLaunchKernelGGL((k_gather<D, NP, NPW, false, false>), grid, block, 0, s, ga);
  }
  PYR_CHECK_LAUNCH();
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
      case 20: return launch<3, 20, 10>(ga, K, a2, count);
      case 35: return launch<3, 35, 12>(ga, K, a2, count);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

// the FP64 matrix-core list gather: MF64_W work items (waves) per workgroup
// (2 measured 2% faster than 1 and 4, 8 slower)
constexpr int MF64_W = 2;
template <int D, int NP>
static void launch_mf64(const ListArgsM64 &la, uint32_t n_items, bool a2) {
  dim3 grid(gather_grid<XCD_CHUNK_MF64>((int)((n_items + MF64_W - 1) / MF64_W))),
      block(64 * MF64_W);
  hipStream_t s = stream();
  if (a2) hipLaunchKernelGGL((k_gather_mf64<D, NP, MF64_W, true>), grid, block, 0, s, la);
  else hipLaunchKernelGGL((k_gather_mf64<D, NP, MF64_W, false>), grid, block, 0, s, la);
  PYR_CHECK_LAUNCH();
}

void launch_gather_mf64(int D, int Np, const ListArgsM64 &la, uint32_t n_items, bool a2) {
  if (D == 2) {
    switch (Np) {
      case 3: return launch_mf64<2, 3>(la, n_items, a2);
      case 6: return launch_mf64<2, 6>(la, n_items, a2);
      case 10: return launch_mf64<2, 10>(la, n_items, a2);
      case 15: return launch_mf64<2, 15>(la, n_items, a2);
    }
  } else {
    switch (Np) {
      case 4: return launch_mf64<3, 4>(la, n_items, a2);
      case 10: return launch_mf64<3, 10>(la, n_items, a2);
      case 20: return launch_mf64<3, 20>(la, n_items, a2);
      case 35: return launch_mf64<3, 35>(la, n_items, a2);
    }
  }
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");
}

}  // namespace pyr
