// gather_fp32.hip -- instantiations of the FP32-compensated deposition over
// culled chunk lists: the cull (k_cull_chunks) and the list gather on the
// matrix cores (k_gather_cl, CL_W work items per workgroup).
#include "deposit_kernels.h"

namespace pyr {

template <int D, int NP>
static void launch_cl(const ListArgs &la, uint32_t n_items, bool a2) {
  dim3 grid(gather_grid((int)((n_items + CL_W - 1) / CL_W))), block(64 * CL_W);
  hipStream_t s = stream();
  if (a2) hipLaunchKernelGGL((k_gather_cl<D, NP, CL_W, true>), grid, block, 0, s, la);
  else hipLaunchKernelGGL((k_gather_cl<D, NP, CL_W, false>), grid, block, 0, s, la);
  PYR_CHECK_LAUNCH();
}

template <int D, int NP>
static void launch_cull(const CullArgs &ca) {
  hipLaunchKernelGGL((k_cull_chunks<D, (NP + 3) / 4>), dim3((unsigned)((ca.K + 3) / 4)),
                     dim3(256), 0, stream(), ca);
  PYR_CHECK_LAUNCH();
}

#define PYR_NP_DISPATCH(FN, ...)                                                          \
  if (D == 2) {                                                                           \
    switch (Np) {                                                                         \
      case 3: return FN<2, 3>(__VA_ARGS__);                                               \
      case 6: return FN<2, 6>(__VA_ARGS__);                                               \
      case 10: return FN<2, 10>(__VA_ARGS__);                                             \
      case 15: return FN<2, 15>(__VA_ARGS__);                                             \
    }                                                                                     \
  } else {                                                                                \
    switch (Np) {                                                                         \
      case 4: return FN<3, 4>(__VA_ARGS__);                                               \
      case 10: return FN<3, 10>(__VA_ARGS__);                                             \
      case 20: return FN<3, 20>(__VA_ARGS__);                                             \
      case 35: return FN<3, 35>(__VA_ARGS__);                                             \
    }                                                                                     \
  }                                                                                       \
  throw Error(PYR_ERR_INVALID_ARG, "deposition supports P1..P4 (2D) and P1..P4 (3D) elements");

void launch_gather_cl(int D, int Np, const ListArgs &la, uint32_t n_items, bool a2) {
  PYR_NP_DISPATCH(launch_cl, la, n_items, a2)
}

void launch_cull_chunks(int D, int Np, const CullArgs &ca) { PYR_NP_DISPATCH(launch_cull, ca) }

#undef PYR_NP_DISPATCH

}  // namespace pyr
