// walk.hip -- the face walk of the reference's shape deposition
// (element_finder.hpp:203-294 face_based_element_finder), run per particle
// for the normalized depositor's statistics (dep_normshape.hpp:131-136
// centroid_distance_stats, :175-177 el_per_particle_stats).  The gather
// deposition does not walk (it reaches every element within R directly), so
// this pass exists only to reproduce those two statistics.
//
// Each thread walks particles in a grid-stride loop.  The depth-first order of
// the reference's recursion is kept with an explicit stack of frames (element,
// next face, periodic offsets, the state of a periodic face's two branches);
// the visited set is a linear-probing hash table in private memory whose slots
// carry a per-particle generation, so it is cleared once per thread.
#include "pyr_internal.h"

namespace pyr {
namespace {

constexpr int WALK_THREADS = 64;
constexpr int WALK_BLOCKS = 4096;
constexpr int WALK_HASH = 512;       // visited-set slots (power of two)
constexpr int WALK_MAX = 384;        // visits per particle before the walk is refused
constexpr int WALK_STACK = WALK_MAX; // recursion depth never exceeds the visits

struct Frame {
  uint32_t en;
  uint32_t meta;   // bits 0-2 next face, 3 second branch pending, 4 first branch
                   // taken, 8-31 periodic offsets (biased by 128) per axis
};

__device__ __forceinline__ int off_of(uint32_t meta, int a) {
  return (int)((meta >> (8 + 8 * a)) & 0xFFu) - 128;
}
__device__ __forceinline__ uint32_t with_off(uint32_t meta, int a, int v) {
  const uint32_t sh = 8 + 8 * a;
  return (meta & ~(0xFFu << sh)) | ((uint32_t)(v + 128) << sh);
}

template <int D>
__global__ __launch_bounds__(WALK_THREADS) void k_face_walk(
    DevMesh m, int64_t n, double R, const double *__restrict__ x,
    const uint32_t *__restrict__ elem, uint32_t *__restrict__ n_el, double *__restrict__ dsum,
    double *__restrict__ dsq, double *__restrict__ dmin, double *__restrict__ dmax,
    int *overflow) {
  uint64_t table[WALK_HASH];
  Frame stack[WALK_STACK];
  for (int k = 0; k < WALK_HASH; ++k) table[k] = ~0ull;
  double L[3] = {0, 0, 0};
  for (int a = 0; a < D; ++a) L[a] = m.per_max[a] - m.per_min[a];
  uint32_t gen = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride, ++gen) {
    const uint32_t e0 = elem[i];
    double xp[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) xp[a] = x[i * D + a];
    uint32_t cnt = 0;
    double s = 0, sq = 0, mn = 0, mx = 0;
    bool over = false;
    auto contains = [&](uint32_t e) {
      uint32_t h = (e * 2654435761u) >> (32 - 9);
      for (;;) {
        const uint64_t v = table[h];
        if ((uint32_t)(v >> 32) != gen) return false;
        if ((uint32_t)v == e) return true;
        h = (h + 1) & (WALK_HASH - 1);
      }
    };
    auto insert = [&](uint32_t e) {
      uint32_t h = (e * 2654435761u) >> (32 - 9);
      for (;;) {
        const uint64_t v = table[h];
        if ((uint32_t)(v >> 32) != gen) {
          table[h] = ((uint64_t)gen << 32) | e;
          return;
        }
        if ((uint32_t)v == e) return;
        h = (h + 1) & (WALK_HASH - 1);
      }
    };
    // add_shape_on_element: one visit of element e with the shape centred at
    // x + offsets * L (dep_normshape.hpp:131-136; meshdata.hpp:169-180 centroid)
    int sp = 0;
    auto visit = [&](uint32_t e, uint32_t meta) {
      if (cnt >= (uint32_t)WALK_MAX) {
        over = true;
        return;
      }
      double c[3] = {0, 0, 0};
      for (int k = 0; k <= D; ++k) {
        const uint32_t v = m.el_vertices[(size_t)e * (D + 1) + k];
        for (int a = 0; a < D; ++a) c[a] += m.vertices[(size_t)v * D + a];
      }
      double d2 = 0;
      for (int a = 0; a < D; ++a) {
        const double dv = c[a] / (D + 1) - (xp[a] + off_of(meta, a) * L[a]);
        d2 += dv * dv;
      }
      const double dist = sqrt(d2);
      if (cnt == 0 || dist < mn) mn = dist;
      if (cnt == 0 || dist > mx) mx = dist;
      ++cnt;
      s += dist;
      sq += dist * dist;
      insert(e);
      stack[sp].en = e;
      stack[sp].meta = meta & ~0xFFu;   // face 0, phase 0
      ++sp;
    };
    if (e0 < (uint32_t)m.K) visit(e0, with_off(with_off(with_off(0u, 0, 0), 1, 0), 2, 0));
    while (sp > 0 && !over) {
      Frame &t = stack[sp - 1];
      const uint32_t f = t.meta & 7u;
      if (f > (uint32_t)D) {
        --sp;
        continue;
      }
      const size_t fi = (size_t)t.en * (D + 1) + f;
      const uint32_t nb = m.face_neighbor[fi];
      double pos[3] = {0, 0, 0};
      for (int a = 0; a < D; ++a) pos[a] = xp[a] + off_of(t.meta, a) * L[a];
      if ((t.meta & 8u) == 0) {
        // the face loop of recurse (element_finder.hpp:246-273)
        t.meta = t.meta + 1;   // next face, unless a periodic face's branches pend
        if (nb == INVALID || contains(nb)) continue;
        double ip = 0;
        for (int a = 0; a < D; ++a) ip += pos[a] * m.face_normal[fi * D + a];
        if (fabs(ip - m.face_rhs[fi]) > R) continue;
        double s2 = 0;
        for (int a = 0; a < D; ++a) {
          const double dv = m.face_centroid[fi * D + a] - pos[a];
          s2 += dv * dv;
        }
        if (sqrt(s2) > R + m.face_radius[fi]) continue;
        const uint32_t axis = m.face_per_axis[fi];
        if (axis == INVALID) {
          visit(nb, t.meta);
          continue;
        }
        // periodic face (:276-290): first branch now, second after it returns
        t.meta = (t.meta - 1) | 8u;
        if (pos[axis] - R < m.per_min[axis]) {
          t.meta |= 16u;
          visit(nb, with_off(t.meta, (int)axis, off_of(t.meta, (int)axis) + 1));
        }
        continue;
      }
      // second branch: pos2 went up by a period if the first branch ran
      const uint32_t axis = m.face_per_axis[fi];
      const int o2 = off_of(t.meta, (int)axis) + ((t.meta & 16u) ? 1 : 0) - 1;
      const uint32_t meta2 = with_off(t.meta, (int)axis, o2);
      t.meta = (t.meta & ~24u) + 1;
      if (pos[axis] + R > m.per_max[axis]) visit(nb, meta2);
    }
    if (over) atomicOr(overflow, 1);
    n_el[i] = cnt;
    dsum[i] = s;
    dsq[i] = sq;
    dmin[i] = mn;
    dmax[i] = mx;
  }
}

}  // namespace
}  // namespace pyr

using namespace pyr;

extern "C" {

int pyr_face_walk_stats(pyr_mesh *mesh, double radius, int64_t n, const double *x,
                        const uint32_t *elem, uint32_t *n_el, double *dist_sum,
                        double *dist_sumsq, double *dist_min, double *dist_max) {
  return guarded([&] {
    PYR_REQUIRE(mesh, "null mesh");
    PYR_REQUIRE(radius > 0 && std::isfinite(radius), "shape radius must be positive");
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    if (n == 0) return;
    PYR_REQUIRE(x && elem && n_el && dist_sum && dist_sumsq && dist_min && dist_max,
                "null pointer");
    mesh->s_misc.reserve(64);
    int *overflow = mesh->s_misc.as<int>();
    hipStream_t s = stream();
    PYR_HIP(hipMemsetAsync(overflow, 0, sizeof(int), s));
    const int64_t need = (n + WALK_THREADS - 1) / WALK_THREADS;
    const unsigned blocks = (unsigned)std::min<int64_t>(WALK_BLOCKS, need);
    if (mesh->dim == 2)
      hipLaunchKernelGGL(k_face_walk<2>, dim3(blocks), dim3(WALK_THREADS), 0, s, mesh->dev, n,
                         radius, x, elem, n_el, dist_sum, dist_sumsq, dist_min, dist_max,
                         overflow);
    else
      hipLaunchKernelGGL(k_face_walk<3>, dim3(blocks), dim3(WALK_THREADS), 0, s, mesh->dev, n,
                         radius, x, elem, n_el, dist_sum, dist_sumsq, dist_min, dist_max,
                         overflow);
    PYR_CHECK_LAUNCH();
    int h = 0;
    PYR_HIP(hipMemcpyAsync(&h, overflow, sizeof(int), hipMemcpyDeviceToHost, s));
    PYR_HIP(hipStreamSynchronize(s));
    if (h)
      throw Error(PYR_ERR_INVALID_ARG,
                  "the face walk of a particle visits more than 384 elements: the shape "
                  "radius is too large for the walk statistics");
  });
}

}  // extern "C"
