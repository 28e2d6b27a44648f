// Internal declarations shared by the HIP translation units of
// libpyrticle_mi355x.so.  Not part of the C ABI (see include/pyrticle_mi355x.h).
//
// Everything here is built with -ffp-contract=off: a*b+c rounds twice unless a
// kernel asks for fma() explicitly.  The localization path relies on that to
// reproduce the reference's FP64 point-in-simplex decisions bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pyrticle_mi355x.h"

namespace pyr {

constexpr uint32_t INVALID = 0xFFFFFFFFu;
constexpr double IN_TOL = 1e-10;   // meshdata.hpp:40 default tolerance

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string &msg);
hipStream_t stream();

#define PYR_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::pyr::Error(PYR_ERR_HIP, std::string(#call) + ": " +                  \
                                          hipGetErrorString(e_));                  \
  } while (0)

#define PYR_CHECK_LAUNCH() PYR_HIP(hipGetLastError())

#define PYR_REQUIRE(cond, msg)                                                     \
  do {                                                                             \
    if (!(cond)) throw ::pyr::Error(PYR_ERR_INVALID_ARG, msg);                     \
  } while (0)

// Optional HIP-event timing of named kernel regions (pyr_profile_enable).
// Events are recorded on the library stream around the region's launches.
bool profiling();
void prof_record(int region, bool begin);
enum {
  PROF_PUSH = 0,
  PROF_GATHER = 1,
  PROF_SORT = 2,
  PROF_COMPACT = 3,
  PROF_ADJOINT = 4,
  PROF_DEPOSIT_ADJ = 5,
  PROF_INTEGRAL = 6,     // shape integrals of the normalized deposition
  PROF_AVERAGE = 7,      // shape-weighted field gather of the averaging pusher
  PROF_DIAG = 8,         // diagnostics reductions
  PROF_NREGIONS = 9
};
struct ProfScope {
  int region;
  explicit ProfScope(int r) : region(r) {
    if (profiling()) prof_record(region, true);
  }
  ~ProfScope() {
    if (profiling()) prof_record(region, false);
  }
};

// Wraps an ABI entry point body: exceptions become error codes.
template <class F>
int guarded(F &&f) {
  try {
    f();
    return PYR_OK;
  } catch (const Error &e) {
    set_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_error(e.what());
    return PYR_ERR_INTERNAL;
  }
}

// Device buffer owned by a handle.  Grows, never shrinks.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t b) {
    if (b <= bytes) return;
    if (ptr) PYR_HIP(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    PYR_HIP(hipMalloc(&ptr, b ? b : 1));
    bytes = b;
  }
  template <class T>
  T *as() const { return static_cast<T *>(ptr); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

template <class T>
T *upload(DevBuf &buf, const T *host, size_t count) {
  buf.reserve(sizeof(T) * count);
  if (count) PYR_HIP(hipMemcpy(buf.ptr, host, sizeof(T) * count, hipMemcpyHostToDevice));
  return buf.as<T>();
}

// Device view of the mesh contract (pyrticle_amd/meshdata.py layout).
struct DevMesh {
  int dim, K, Np, nv;
  const double *inv_map;        // K x (d*d + d)
  const double *jacobian;       // K: |det| of the forward maps (or null)
  const uint32_t *el_vertices;  // K x (d+1)
  const double *face_normal;    // K x (d+1) x d
  const double *face_rhs;       // K x (d+1)
  const double *face_centroid;  // K x (d+1) x d
  const double *face_radius;    // K x (d+1)
  const uint32_t *face_neighbor;
  const uint32_t *face_per_axis;
  const double *vertices;       // nv x d
  const uint32_t *vadj_starts;  // nv + 1
  const uint32_t *vadj_elements;
  const double *nodes;          // K*Np x d
  double per_min[3], per_max[3];
  // uniform grid over the (slightly inflated) mesh bounding box for the
  // global-search fallback; cell lists are sorted by element id.
  double g_lo[3], g_inv_h[3];
  int g_n[3];
  const uint32_t *g_starts;
  const uint32_t *g_elems;
};

// ---------------------------------------------------------------------------
// localization, restating particle_state.hpp:130-261 and meshdata.hpp:38-48.
// ---------------------------------------------------------------------------
struct FindCounters {
  uint32_t same = 0, neighbor = 0, vertex = 0, global = 0;
};

template <int D>
__device__ __forceinline__ void inverse_map(const DevMesh &m, uint32_t en, const double *x,
                                            double *r) {
  const double *a = m.inv_map + (size_t)en * (D * D + D);
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double t = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) t += a[i * D + j] * x[j];
    r[i] = t + a[D * D + i];
  }
}

template <int D>
__device__ __forceinline__ bool in_unit_simplex(const double *r) {
#pragma unroll
  for (int i = 0; i < D; ++i)
    if (r[i] < -1 - IN_TOL) return false;
  double s = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) s += r[i];
  return s <= (double)(-(D - 2)) + IN_TOL;
}

template <int D>
__device__ __forceinline__ bool in_element(const DevMesh &m, uint32_t en, const double *x,
                                           double *r) {
  inverse_map<D>(m, en, x, r);
  return in_unit_simplex<D>(r);
}

// meshdata.hpp:208-215: the lowest element index containing x.  The grid cell
// of x lists every element whose inflated bounding box meets the cell, in
// ascending order, so the first hit is the reference's answer.
template <int D>
__device__ uint32_t find_global(const DevMesh &m, const double *x, double *r) {
  int cell = 0;
#pragma unroll
  for (int i = D - 1; i >= 0; --i) {
    double c = (x[i] - m.g_lo[i]) * m.g_inv_h[i];
    if (!(c >= 0.0) || c >= (double)m.g_n[i]) return INVALID;
    cell = cell * m.g_n[i] + (int)c;
  }
  for (uint32_t k = m.g_starts[cell]; k < m.g_starts[cell + 1]; ++k) {
    uint32_t en = m.g_elems[k];
    if (in_element<D>(m, en, x, r)) return en;
  }
  return INVALID;
}

// The tail of the cascade (particle_state.hpp:198-261): the elements around
// the closest vertex of prev, then the global search.  (Out of line it saved
// three registers and cost the push 2x in spills around the call.)
template <int D>
__device__ __forceinline__ uint32_t find_by_vertex_or_global(const DevMesh &m,
                                                             const double *x, uint32_t prev,
                                                             FindCounters &cnt, double *r) {
  if (prev != INVALID && prev < (uint32_t)m.K) {
    uint32_t closest_vertex = INVALID;
    double min_dist = __builtin_inf();
#pragma unroll
    for (int k = 0; k <= D; ++k) {
      uint32_t vi = m.el_vertices[(size_t)prev * (D + 1) + k];
      double s = 0;
#pragma unroll
      for (int i = 0; i < D; ++i) {
        double dv = m.vertices[(size_t)vi * D + i] - x[i];
        s += dv * dv;
      }
      double dist = sqrt(s);
      if (dist < min_dist) {
        closest_vertex = vi;
        min_dist = dist;
      }
    }
    for (uint32_t k = m.vadj_starts[closest_vertex]; k < m.vadj_starts[closest_vertex + 1];
         ++k) {
      uint32_t possible2 = m.vadj_elements[k];
      if (in_element<D>(m, possible2, x, r)) {
        cnt.vertex++;
        return possible2;
      }
    }
  }
  cnt.global++;
  return find_global<D>(m, x, r);
}

// particle_state.hpp:160-261: the cascade once x is known to lie outside prev
// (the neighbour across the face whose normal best matches the momentum, then
// the closest-vertex fan, then the global search).
template <int D>
__device__ uint32_t find_outside_prev(const DevMesh &m, const double *x, const double *p,
                                      uint32_t prev, FindCounters &cnt, int &err, double *r) {
  if (prev != INVALID && prev < (uint32_t)m.K) {
    int closest = -1;
    double max_ip = 0;
#pragma unroll
    for (int f = 0; f <= D; ++f) {
      const double *n = m.face_normal + ((size_t)prev * (D + 1) + f) * D;
      double ip = 0;
#pragma unroll
      for (int i = 0; i < D; ++i) ip += n[i] * p[i];
      if (ip > max_ip) {
        closest = f;
        max_ip = ip;
      }
    }
    if (closest == -1) {
      err = 1;
      return INVALID;
    }
    uint32_t possible = m.face_neighbor[(size_t)prev * (D + 1) + closest];
    if (possible != INVALID && in_element<D>(m, possible, x, r)) {
      cnt.neighbor++;
      return possible;
    }
  }
  return find_by_vertex_or_global<D>(m, x, prev, cnt, r);
}

// particle_state.hpp:130-261.  *err is set when the reference would throw
// "no best normal found--weird".  r receives the unit coordinates of x in the
// returned element (valid when the result is not INVALID).
template <int D>
__device__ uint32_t find_new_containing_element(const DevMesh &m, const double *x,
                                                const double *p, uint32_t prev,
                                                FindCounters &cnt, int &err, double *r) {
  if (prev != INVALID && prev < (uint32_t)m.K && in_element<D>(m, prev, x, r)) {
    cnt.same++;
    return prev;
  }
  return find_outside_prev<D>(m, x, p, prev, cnt, err, r);
}

// boundary_hit for a particle no element contains (particle_state.hpp:
// 301-357): on a periodic trip x is wrapped in place, as the reference writes
// it back (:341), and located again.  Returns false if the particle dies.
template <int D>
__device__ __forceinline__ bool boundary_hit(const DevMesh &m, double *x, const double *p,
                                             uint32_t &elem, FindCounters &cnt, int &err,
                                             double *r) {
  bool trip = false;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double mn = m.per_min[a], mx = m.per_max[a];
    if (mn != mx) {
      const double xi = x[a];
      if (xi < mn) {
        x[a] += (mx - mn);
        trip = true;
      } else if (xi > mx) {
        x[a] -= (mx - mn);
        trip = true;
      }
    }
  }
  if (trip) {
    uint32_t ce = find_new_containing_element<D>(m, x, p, elem, cnt, err, r);
    if (err) return false;
    if (ce != INVALID) {
      // the reference re-examines the particle at the same index, which
      // counts one more "same" hit (update_containing_elements loop, :274-294)
      cnt.same++;
      elem = ce;
      return true;
    }
  }
  return false;
}

// Read-only data seen through the constant address space: loads with a
// wave-uniform address become scalar loads into SGPRs.
template <class T>
using const_ptr = const T __attribute__((address_space(4))) *;
template <class T>
__device__ __forceinline__ const_ptr<T> as_constant(const T *p) {
  return (const_ptr<T>)p;
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// First element of list[k0, k1) (in list order) that contains x, tested by
// the wave's active lanes 64 at a time; INVALID if none.  rl receives the
// unit coordinates in the element found (from the lane that tested it).
template <int D>
__device__ __forceinline__ uint32_t wave_first_containing(const DevMesh &m,
                                                          const uint32_t *list, uint32_t k0,
                                                          uint32_t k1, const double *x,
                                                          double *rl) {
  const uint64_t act = __ballot(1);
  const int nact = __popcll(act);
  const int rank = __popcll(act & ((1ull << (threadIdx.x & 63)) - 1));
  for (uint32_t b = k0; b < k1; b += (uint32_t)nact) {
    const uint32_t k = b + (uint32_t)rank;
    uint32_t cand = INVALID;
    double rr[D];
    bool hit = false;
    if (k < k1) {
      cand = list[k];
      hit = in_element<D>(m, cand, x, rr);
    }
    const uint64_t hb = __ballot(hit);
    if (hb) {
      const int j = (int)__builtin_ctzll(hb);   // lowest hit lane = lowest k
#pragma unroll
      for (int i = 0; i < D; ++i) rl[i] = readlane_f64(rr[i], j);
      return (uint32_t)__builtin_amdgcn_readlane((int)cand, j);
    }
  }
  return INVALID;
}

// find_by_vertex_or_global (particle_state.hpp:198-261, meshdata.hpp:208-215)
// for the particle of lane L (position x, previous element prev: uniform
// values), run by the whole wave: the elements around prev's closest vertex
// (when prev is an element),
// then the global grid cell, each list tested 64 elements at a time and the
// first containing one (in list order) taken, as the sequential scan would.
// kind: 1 vertex fan, 2 global search (the counter the lane bumps).
template <int D>
__device__ __forceinline__ uint32_t wave_vertex_or_global(const DevMesh &m, const double *x,
                                                          uint32_t prev, double *rl,
                                                          int &kind) {
  if (prev != INVALID && prev < (uint32_t)m.K) {
  uint32_t closest_vertex = INVALID;
  double min_dist = __builtin_inf();
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    const uint32_t vi = as_constant(m.el_vertices)[(size_t)prev * (D + 1) + k];
    double s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const double dv = as_constant(m.vertices)[(size_t)vi * D + i] - x[i];
      s += dv * dv;
    }
    const double dist = sqrt(s);
    if (dist < min_dist) {
      closest_vertex = vi;
      min_dist = dist;
    }
  }
  closest_vertex = (uint32_t)__builtin_amdgcn_readfirstlane((int)closest_vertex);
  uint32_t res = wave_first_containing<D>(m, m.vadj_elements,
                                          as_constant(m.vadj_starts)[closest_vertex],
                                          as_constant(m.vadj_starts)[closest_vertex + 1], x, rl);
  if (res != INVALID) {
    kind = 1;
    return res;
  }
  }
  kind = 2;
  int cell = 0;
#pragma unroll
  for (int i = D - 1; i >= 0; --i) {
    const double c = (x[i] - m.g_lo[i]) * m.g_inv_h[i];
    if (!(c >= 0.0) || c >= (double)m.g_n[i]) return INVALID;
    cell = cell * m.g_n[i] + (int)c;
  }
  cell = __builtin_amdgcn_readfirstlane(cell);
  return wave_first_containing<D>(m, m.g_elems, as_constant(m.g_starts)[cell],
                                  as_constant(m.g_starts)[cell + 1], x, rl);
}

// The tail of the cascade for the active lanes with head == 2, one particle
// at a time with the whole wave (a lane's own sequential fan costs its wave two
// dependent loads per candidate element).  A lane whose particle some element
// contains gets head = 3 with elem and r set; head stays 2 when none does.
template <int D>
__device__ __forceinline__ void wave_cascade_tail(const DevMesh &m, const double *x,
                                                  uint32_t &elem, FindCounters &cnt, double *r,
                                                  int &head) {
  uint64_t pend = __ballot(head == 2);
  while (pend) {
    const int L = (int)__builtin_ctzll(pend);
    pend &= pend - 1;
    double xl[D], rl[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xl[k] = readlane_f64(x[k], L);
    const uint32_t prev = (uint32_t)__builtin_amdgcn_readlane((int)elem, L);
    int kind = 0;
    const uint32_t ne = wave_vertex_or_global<D>(m, xl, prev, rl, kind);
    if ((int)(threadIdx.x & 63) == L) {
      if (kind == 1) cnt.vertex++;
      else cnt.global++;
      if (ne != INVALID) {
        elem = ne;
#pragma unroll
        for (int k = 0; k < D; ++k) r[k] = rl[k];
        head = 3;
      }
    }
  }
}

// update_containing_elements + boundary_hit (particle_state.hpp:266-357) for
// the active lanes with need: the cascade of find_new_containing_element from
// the element held (skipping its first test when outside: the caller found x
// outside that element with the same arithmetic), then the periodic trip or
// death.  Returns false if the particle dies; the other lanes keep their
// element and return true.  The fan and grid searches are shared by the wave,
// so call it with every lane that can help active, not from a branch on need.
template <int D>
__device__ __forceinline__ bool relocate_wave(const DevMesh &m, bool need, bool outside,
                                              double *x, const double *p, uint32_t &elem,
                                              FindCounters &cnt, int &err, double *r) {
  int head = 1;
  if (need) {
    head = 2;
    const bool valid = elem != INVALID && elem < (uint32_t)m.K;
    if (valid && !outside && in_element<D>(m, elem, x, r)) {
      cnt.same++;
      head = 1;
    } else if (valid) {
      int closest = -1;
      double max_ip = 0;
#pragma unroll
      for (int f = 0; f <= D; ++f) {
        const double *n = m.face_normal + ((size_t)elem * (D + 1) + f) * D;
        double ip = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) ip += n[i] * p[i];
        if (ip > max_ip) {
          closest = f;
          max_ip = ip;
        }
      }
      if (closest == -1) {
        err = 1;
        head = 0;
      } else {
        const uint32_t possible = m.face_neighbor[(size_t)elem * (D + 1) + closest];
        if (possible != INVALID && in_element<D>(m, possible, x, r)) {
          cnt.neighbor++;
          elem = possible;
          head = 1;
        }
      }
    }
  }
  wave_cascade_tail<D>(m, x, elem, cnt, r, head);
  if (head == 0) return false;
  if (head == 2) return boundary_hit<D>(m, x, p, elem, cnt, err, r);
  return true;
}

// particle_state.hpp:98-125 get_velocities, same operation order.
template <int VD>
__device__ __forceinline__ void velocity(const double *p, double m, double c, double *v) {
  double s = 0;
#pragma unroll
  for (int i = 0; i < VD; ++i) s += p[i] * p[i];
  double pm = sqrt(s);
  if (pm == 0) {
#pragma unroll
    for (int i = 0; i < VD; ++i) v[i] = 0;
  } else {
    double vm = c * pm / sqrt(m * m * c * c + pm * pm);
    double f = vm / pm;
#pragma unroll
    for (int i = 0; i < VD; ++i) v[i] = f * p[i];
  }
}

// Wave-aggregated accumulation of the four find counters.
__device__ __forceinline__ void flush_counters(const FindCounters &c,
                                               unsigned long long *out) {
  if (!out) return;
  unsigned long long vals[4] = {c.same, c.neighbor, c.vertex, c.global};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    unsigned long long v = vals[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(out + k, v);
  }
}

// Exclusive scan of n uint32 values (device pointers, stream-ordered).
// `total` (device, may be null) receives the sum.  tmp must hold
// scan_tmp_words(n) words.
size_t scan_tmp_words(int64_t n);
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total,
                        uint32_t *tmp, hipStream_t s);

}  // namespace pyr

// Orders a wave's LDS writes before its later LDS reads of other lanes' data
// (the compiler sees no dependence between them).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

namespace pyr {
// Block b of a grid of nblocks -> the block index it handles, such that chunks
// of 128 consecutive indices go to one XCD (the hardware deals consecutive
// blocks round-robin to the 8 XCDs).  Kernels that read particles in a
// shuffled order (the staging after the element sort, the push in that order)
// then fetch each element's particles through one L2 instead of all eight
// (runs of 128 against 32: steps 0.7% shorter on the 4M-particle cylinder,
// runs of 8 no faster than 32).  A bijection on [0, nblocks): the blocks past
// the last whole group of 8 chunks map to themselves.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
  constexpr int64_t C = 128;
  const int64_t full = nblocks / (8 * C) * (8 * C);
  if (b >= full) return b;
  const int64_t xcd = b & 7, slot = b >> 3;
  return ((slot / C) * 8 + xcd) * C + slot % C;
}
}  // namespace pyr

// Opaque handle behind the ABI's pyr_mesh.
struct pyr_mesh;
namespace pyr {
// deposit.hip: the permutation (sorted position -> particle index) left by the
// last sort of particles by element on this mesh, if it sorted n particles,
// else null.  Any permutation of [0, n) is a
// valid processing order for the fused push, whose particles are independent;
// the last deposition's sort groups the particles of each element into
// consecutive lanes, so a wave mostly sees one element.
const uint32_t *deposit_last_order(pyr_mesh *mesh, int64_t n);
// deposit.hip: a stable sort of n particles by element (elem, device): the
// permutation sorted position -> particle index, valid until the next sort on
// this mesh (stream-ordered).
const uint32_t *element_order(pyr_mesh *mesh, int64_t n, const uint32_t *elem);
// core.hip: copies the two ints at dev (stream-ordered) into out, through
// the mesh's pinned words and an event, and waits for them (not for work
// queued later).
void read_step_flags(pyr_mesh *mesh, const int *dev, int out[2]);
// sort.hip: stable radix sort of the indices 0..n-1 by keys clamped to
// [0, kmax] (stream-ordered on s): sorted keys in kout, sorted indices in
// vout (returned); scratch grows to the sort's temporary storage.
const uint32_t *sort_pairs(DevBuf &scratch, int64_t n, const uint32_t *keys, uint32_t kmax,
                           uint32_t *kout, uint32_t *vout, hipStream_t s);
// deposit.hip: the averaging pusher's shape-weighted forces, stream-ordered
// (returns the device count of particles with a zero shape integral, or null
// when n = 0).  vis: six arrays (pt_e, pt_b, pt_e_stddev, pt_b_stddev,
// el_force, mag_force) or null.
unsigned long long *average_forces_impl(pyr_mesh *mesh, double radius, double alpha,
                                        double normalizer, int64_t n, int vdim,
                                        const double *const fields[6], const double *x,
                                        const uint32_t *elem, const double *q, const double *v,
                                        const double *integral_weights, double *force,
                                        double *integral_out, double *const *vis);
}  // namespace pyr

struct pyr_mesh {
  pyr::DevMesh dev{};
  int dim = 0, K = 0, Np = 0, nv = 0;
  std::vector<pyr::DevBuf> owned;   // mesh arrays
  // interpolation: modal transform Minv = V^{-1} (Np x Np, row-major), exps
  pyr::DevBuf minv, minv32, exps;
  std::vector<int32_t> h_exps;
  // scratch
  pyr::DevBuf s_modal, s_flags, s_scan, s_tmp, s_holes, s_misc;
  pyr::DevBuf s_adj;                // adjoint recompute buffers (push.hip)
  // deposition state (see deposit.hip)
  struct DepositCache *dep = nullptr;
  pyr::DevBuf counters;             // 4 x u64 scratch for counters
  // static fields as modal coefficients, K x 6 x Np (pyr_set_static_fields)
  pyr::DevBuf sf64, sf32;
  bool sf_set = false;
  // the same fields nodal, 6 x K x Np, for the averaging pusher's step;
  // sf_has[f]: field f was given (else zero)
  pyr::DevBuf sf_nodal;
  bool sf_has[6] = {false, false, false, false, false, false};
  pyr::DevBuf s_avg;                // averaging step scratch (push.hip)
  // pyr_forces: occ[e] == occ_stamp marks element e occupied in this call
  pyr::DevBuf s_occ;
  uint32_t occ_stamp = 0;
  // measurement switches of this mesh (no reference counterpart):
  // push_mfma: FP32C push interpolation on the matrix cores (0 off, 1 on,
  // 2 on and count the particle-stages that took it in mf_count);
  // deposit_count: pyr_deposit also counts pair and visit statistics
  int push_mfma = 0;
  pyr::DevBuf mf_count;
  int deposit_count = 0;
  // pinned words and event of the fused steps' per-step flag read-back
  int *h_step = nullptr;
  hipEvent_t ev_step = nullptr;
  ~pyr_mesh();
};
