// deposit.hip -- shape-function deposition of rho and j onto DG element nodes.
//
// The reference scatters: for every particle it walks the face-connected
// elements its shape touches and adds q S(node - x) to each node
// (dep_shape.hpp:128-149, element_finder.hpp:203-294, dep_target.hpp:54-186).
// Here the same sum is gathered per target element, with no float atomics:
//
//   1. particles are sorted by containing element with a stable LSD radix
//      sort (8-bit digits), giving contiguous per-element segments;
//   2. each target element owns a candidate list of source elements that can
//      hold a particle whose shape reaches it (built on the host once per
//      radius from bounding-sphere and face-plane separation tests);
//   3. a block per target element streams the candidate particles, drops
//      those whose ball cannot reach its nodes, queues the rest in LDS and
//      evaluates them one particle per lane against the element's nodes, held
//      in scalar registers, accumulating rho and j per node in registers.
//
// Summation order is fixed by the candidate order and the stable sort, so
// results are bitwise reproducible run to run.  In a convex conforming mesh the
// nodes within R of a particle are exactly the nodes of elements its ball
// meets, so the gather and the reference's face walk sum the same terms.
#include <algorithm>
#include <array>
#include <tuple>
#include <atomic>
#include <cstdlib>
#include <cmath>
#include <thread>

#include "deposit_kernels.h"

namespace pyr {

// ---------------------------------------------------------------------------
// stable LSD radix sort of (key, index) pairs, 8-bit digits
// ---------------------------------------------------------------------------
constexpr int RS_BS = 256, RS_ROUNDS = 16, RS_TILE = RS_BS * RS_ROUNDS;

__global__ __launch_bounds__(RS_BS) void k_rs_hist(const uint32_t *__restrict__ keys, int64_t n,
                                                   int shift, uint32_t *__restrict__ hist,
                                                   int ntiles, uint32_t kmax) {
  __shared__ uint32_t h[256];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ROUNDS; ++k) {
    int64_t idx = base + (int64_t)k * RS_BS + tid;
    if (idx < n) atomicAdd(&h[(min(keys[idx], kmax) >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)tid * ntiles + blockIdx.x] = h[tid];
}

__global__ __launch_bounds__(RS_BS) void k_rs_scatter(const uint32_t *__restrict__ kin,
                                                      const uint32_t *__restrict__ vin,
                                                      uint32_t *__restrict__ kout,
                                                      uint32_t *__restrict__ vout, int64_t n,
                                                      int shift, const uint32_t *__restrict__ offs,
                                                      int ntiles, uint32_t kmax) {
  __shared__ uint32_t base_pos[256];
  __shared__ uint32_t wcnt[RS_BS / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  base_pos[tid] = offs[(size_t)tid * ntiles + blockIdx.x];
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ROUNDS; ++k) {
    const int64_t idx = base + (int64_t)k * RS_BS + tid;
    const bool valid = idx < n;
    const uint32_t key = valid ? min(kin[idx], kmax) : 0u;
    const uint32_t d = (key >> shift) & 255u;
#pragma unroll
    for (int w = 0; w < RS_BS / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base_pos[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      kout[pos] = key;
      vout[pos] = vin ? vin[idx] : (uint32_t)idx;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < RS_BS / 64; ++w) add += wcnt[w][tid];
    base_pos[tid] += add;
  }
}

// Segment starts from the sorted (clamped) keys: seg[e] = first index whose
// key is >= e << shift, for e = 0..K (seg[K] = number of located particles;
// keys of particles without an element are clamped to K << shift and sort
// last).  One binary search per element, balanced whatever the occupancy.
__global__ void k_seg_from_sorted(int K, int64_t n, const uint32_t *__restrict__ keys,
                                  int shift, uint32_t *__restrict__ seg) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e > K) return;
  const uint32_t target = (uint32_t)e << shift;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  seg[e] = (uint32_t)lo;
}

// staged particles in element-sorted order, two arrays of double4:
//   pos[j] = (x0, x1, x2, 0)          read for every candidate visit
//   wgt[j] = (a, a*v0, a*v1, a*v2)    read only for particles that pass,
// with a = q * normalizer (the charge normalization stays FP64).
__global__ void k_stage(int D, int64_t n, const uint32_t *__restrict__ perm,
                        const double *__restrict__ x, const double *__restrict__ q,
                        const double *__restrict__ v, double normalizer,
                        double4 *__restrict__ pos, double4 *__restrict__ wgt) {
  const int64_t j = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  double xs[3], bs[3];
  const double a = q[i] * normalizer;
  for (int k = 0; k < 3; ++k) {
    xs[k] = k < D ? x[(size_t)i * D + k] : 0.0;
    bs[k] = (v && k < D) ? a * v[(size_t)i * D + k] : 0.0;
  }
  pos[j] = make_double4(xs[0], xs[1], xs[2], 0.0);
  wgt[j] = make_double4(a, bs[0], bs[1], bs[2]);
}







// sort keys for the deposition: element id, then a 12-bit Morton code of the
// position in a 16^D grid over the element's vertex-bounding cube, so that
// consecutive particles of an element are spatial neighbours.  Invalid
// elements get the largest key.
constexpr int MORTON_BITS = 12;
template <int D>
__global__ void k_sort_keys(int64_t n, int K, const uint32_t *__restrict__ elem,
                            const double *__restrict__ x, const double *__restrict__ geom,
                            uint32_t *__restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t e = elem[i];
  if (e >= (uint32_t)K) {
    keys[i] = (uint32_t)K << MORTON_BITS;
    return;
  }
  constexpr int BPA = MORTON_BITS / D;   // bits per axis
  const double vr = geom[(size_t)e * 8 + 4];
  const double scale = vr > 0 ? (double)(1 << BPA) / (2.0 * vr) : 0.0;
  uint32_t code = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double u = (x[i * D + a] - geom[(size_t)e * 8 + a] + vr) * scale;
    const int qv = (int)fmin(fmax(u, 0.0), (double)((1 << BPA) - 1));
#pragma unroll
    for (int b = 0; b < BPA; ++b) code |= (uint32_t)((qv >> b) & 1) << (b * D + a);
  }
  keys[i] = (e << MORTON_BITS) | code;
}

// node table: n - c_e and |n - c_e|^2 / R, per target element
template <class T>
__global__ void k_node_table(int D, int K, int np, const double *__restrict__ nodes,
                             const double *__restrict__ geom, double invR,
                             T *__restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)K * np) return;
  const int e = (int)(t / np);
  double d[3] = {0, 0, 0}, s = 0;
  for (int k = 0; k < D; ++k) {
    d[k] = nodes[(size_t)t * D + k] - geom[(size_t)e * 8 + k];
    s += d[k] * d[k];
  }
  out[(size_t)t * 4 + 0] = (T)d[0];
  out[(size_t)t * 4 + 1] = (T)d[1];
  out[(size_t)t * 4 + 2] = (T)d[2];
  out[(size_t)t * 4 + 3] = (T)(s * invR);
}


// Grouped node tables of the list gathers: slot 4 g + i of element e holds
// node node_slot[4 g + i], in FP32 as (2 (n - c_e) / (R P), -|n - c_e|^2 /
// (R P)) (padding slots (0, 0, 0, -1): t = c_p - 1 <= 0) and in FP64 as
// (n - c_e, |n - c_e|^2 / R), the values of k_node_table (padding slots
// (0, 0, 0, 2 R): t <= -R).
__global__ void k_node_table_g(int D, int K, int np, int ng, const double *__restrict__ nodes,
                               const double *__restrict__ geom, double invR, double inv_p,
                               const int *__restrict__ node_slot, float4 *__restrict__ tab,
                               double4 *__restrict__ tab64) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)K * ng) return;
  const int e = (int)(t / ng), grp = (int)(t % ng);
  for (int i = 0; i < 4; ++i) {
    const int node = node_slot[4 * grp + i];
    const size_t slot = (size_t)e * 4 * ng + 4 * grp + i;
    if (node < 0) {
      tab[slot] = make_float4(0.f, 0.f, 0.f, -1.f);
      // FP64 gather: t = (R - |d|^2 / R) - 2 R <= -R, a zero shape value
      tab64[slot] = make_double4(0.0, 0.0, 0.0, 2.0 / invR);
      continue;
    }
    double d[3], s2 = 0;
    for (int k = 0; k < 3; ++k) {
      d[k] = k < D ? nodes[((size_t)e * np + node) * D + k] - geom[(size_t)e * 8 + k] : 0.0;
      s2 += d[k] * d[k];
    }
    tab64[slot] = make_double4(d[0], d[1], d[2], s2 * invR);
    const double bs = 2.0 * invR * inv_p;   // B = 2 (n - c_e) / (R P)
    tab[slot] = make_float4((float)(d[0] * bs), (float)(d[1] * bs), (float)(d[2] * bs),
                            (float)(-(s2 * invR) * inv_p));
  }
}

// Chunk staging of the cluster gather: chunk c (one wave) = the particles
// [first, first + count) of the element-sorted order, all in source element
// s, in slots 64 c .. 64 c + 63:
//   pos[64 c + l]  = (x - c_s, 0) in FP32 (formed in FP64, rounded once),
//   wgt[64 c + 4 (l / 4) + k][l % 4] = w_l[k], w = (a, a v) in FP32, a = q *
//                    scale formed in FP64 (the weights of each quad of slots,
//                    transposed),
// padding slots (l >= count) at the source centroid with zero weights; and
// meta[c] = a sphere around the FP32 offsets of the chunk (centre of their
// bounding box, radius to the farthest, rounded up).
__global__ __launch_bounds__(256) void k_stage32c(int D, int64_t nch,
                                                  const uint2 *__restrict__ chunks,
                                                  const uint32_t *__restrict__ seg,
                                                  const uint32_t *__restrict__ perm,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ v, double scale,
                                                  const double *__restrict__ geom,
                                                  float4 *__restrict__ pos,
                                                  float *__restrict__ wgt_t,
                                                  float4 *__restrict__ meta) {
  const int lane = threadIdx.x & 63;
  // XCD-chunked blocks: the chunks of one element (consecutive) read the same
  // lines of x, q and v through one L2
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (c >= nch) return;   // uniform per wave
  const uint2 ch = chunks[c];
  const uint32_t s = ch.x, first = ch.y;
  const uint32_t count = min(64u, seg[s + 1] - first);
  const bool valid = (uint32_t)lane < count;
  float xs[3] = {0.f, 0.f, 0.f}, w[4] = {0.f, 0.f, 0.f, 0.f};
  if (valid) {
    const uint32_t i = perm[first + lane];
    const double a = q[i] * scale;
    w[0] = (float)a;
    for (int k = 0; k < D; ++k) {
      xs[k] = (float)(x[(size_t)i * D + k] - geom[(size_t)s * 8 + k]);
      if (v) w[1 + k] = (float)(a * v[(size_t)i * D + k]);
    }
  }
  const int64_t slot = 64 * c + lane;
  pos[slot] = make_float4(xs[0], xs[1], xs[2], 0.f);
  {
    // the quad's weights transposed through the wave's own LDS row: one
    // 16-byte store per lane (component k of the quad's four particles)
    __shared__ float4 s_w[4][64];
    const int wv = threadIdx.x >> 6;
    s_w[wv][lane] = make_float4(w[0], w[1], w[2], w[3]);
    wave_lds_fence();
    const int qb = lane & ~3, k = lane & 3;
    const float *src = reinterpret_cast<const float *>(&s_w[wv][qb]);
    reinterpret_cast<float4 *>(wgt_t)[slot] =
        make_float4(src[0 * 4 + k], src[1 * 4 + k], src[2 * 4 + k], src[3 * 4 + k]);
  }
  float mn[3], mx[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = valid ? xs[k] : INFINITY;
    mx[k] = valid ? xs[k] : -INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
    }
  float ctr[3];
  for (int k = 0; k < 3; ++k) ctr[k] = 0.5f * (mn[k] + mx[k]);
  float r2 = 0.f;
  if (valid)
    for (int k = 0; k < 3; ++k) r2 = fmaf(xs[k] - ctr[k], xs[k] - ctr[k], r2);
  for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o, 64));
  if (lane == 0) meta[c] = make_float4(ctr[0], ctr[1], ctr[2], sqrtf(r2) * (1.f + 1e-5f) + 1e-30f);
}

// Chunk staging of the FP64 matrix-core gather, chunk c (one wave) = the
// particles [first, first + count) of the element-sorted order, all in source
// element s, four slices of 16 (slot 16 sl + r):
//   pos_t[256 c + 64 sl + l] = row l / 16 of the slice's particle l % 16:
//                              p = x - c_s (3), then -|p|^2 / R;
//   wgt_t[256 c + 64 sl + l] = component l % 4 of particle
//                              4 ((l / 4) % 4) + l / 16: (a, a v), a = q scale;
// zeros for padding slots (slot >= count); and meta[c], the chunk's sphere of
// k_stage32c (from the same FP32-rounded offsets, so both precisions cull the
// same lists).  The rows pass through a wave-private LDS transpose.
__global__ __launch_bounds__(256) void k_stage64c(int D, int64_t nch,
                                                  const uint2 *__restrict__ chunks,
                                                  const uint32_t *__restrict__ seg,
                                                  const uint32_t *__restrict__ perm,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ v, double scale,
                                                  const double *__restrict__ geom, double invR,
                                                  double *__restrict__ pos_t,
                                                  double *__restrict__ wgt_t,
                                                  float4 *__restrict__ meta) {
  __shared__ double s_p[4][64][4], s_w[4][64][4];   // per wave: rows of its 64 slots
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t c = xcd_block(blockIdx.x, gridDim.x) * 4 + wv;
  if (c >= nch) return;   // uniform per wave (the LDS rows are the wave's own)
  const uint2 ch = chunks[c];
  const uint32_t s = ch.x, first = ch.y;
  const uint32_t count = min(64u, seg[s + 1] - first);
  const bool valid = (uint32_t)lane < count;
  double p[3] = {0.0, 0.0, 0.0}, w[4] = {0.0, 0.0, 0.0, 0.0};
  float xs[3] = {0.f, 0.f, 0.f};
  if (valid) {
    const uint32_t i = perm[first + lane];
    const double a = q[i] * scale;
    w[0] = a;
    for (int k = 0; k < D; ++k) {
      p[k] = x[(size_t)i * D + k] - geom[(size_t)s * 8 + k];
      xs[k] = (float)p[k];
      if (v) w[1 + k] = a * v[(size_t)i * D + k];
    }
  }
  s_p[wv][lane][0] = p[0];
  s_p[wv][lane][1] = p[1];
  s_p[wv][lane][2] = p[2];
  s_p[wv][lane][3] = -fma(p[2], p[2], fma(p[1], p[1], p[0] * p[0])) * invR;
  for (int k = 0; k < 4; ++k) s_w[wv][lane][k] = w[k];
  wave_lds_fence();
  for (int sl = 0; sl < 4; ++sl) {
    pos_t[(size_t)c * 256 + 64 * sl + lane] = s_p[wv][16 * sl + (lane & 15)][lane >> 4];
    wgt_t[(size_t)c * 256 + 64 * sl + lane] =
        s_w[wv][16 * sl + 4 * ((lane >> 2) & 3) + (lane >> 4)][lane & 3];
  }
  float mn[3], mx[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = valid ? xs[k] : INFINITY;
    mx[k] = valid ? xs[k] : -INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], o, 64));
    }
  float ctr[3];
  for (int k = 0; k < 3; ++k) ctr[k] = 0.5f * (mn[k] + mx[k]);
  float r2 = 0.f;
  if (valid)
    for (int k = 0; k < 3; ++k) r2 = fmaf(xs[k] - ctr[k], xs[k] - ctr[k], r2);
  for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o, 64));
  if (lane == 0) meta[c] = make_float4(ctr[0], ctr[1], ctr[2], sqrtf(r2) * (1.f + 1e-5f) + 1e-30f);
}

// rho and j of every target from its work items' partial sums, added in item
// order (FP64): one thread per (target, node slot).
// Statistics of the chunk lists of the last FP32C deposition, per target:
// list entries, listed (chunk, node group) pairs, and the 4x4x1 matrix-core
// instructions the list gather issues for them (entries 2k and 2k + 1 of a
// list form one pair, evaluated against the union of their groups with D + 6
// products per chunk and group: D + 2 for t, four for the accumulation).
__global__ void k_list_stats(int K, int D, const uint32_t *__restrict__ cap_off,
                             const uint32_t *__restrict__ n_live, const uint2 *__restrict__ ent,
                             unsigned long long *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= K) return;
  const uint32_t b = cap_off[e], n = n_live[e];
  if (!n) return;
  unsigned long long groups = 0, instr = 0;
  for (uint32_t i = 0; i < n; i += 2) {
    const uint32_t ma = ent[b + i].y & ENT_MASK,
                   mb = i + 1 < n ? ent[b + i + 1].y & ENT_MASK : 0u;
    groups += (unsigned)(__popc(ma) + __popc(mb));
    instr += 2ull * (unsigned)(D + 6) * (unsigned)__popc(ma | mb);
  }
  atomicAdd(out + 0, (unsigned long long)n);
  atomicAdd(out + 1, groups);
  atomicAdd(out + 2, instr);
}

struct CombineArgs {
  int K, NS;                   // targets, node slots (4 NG)
  int Np, D;
  const uint32_t *order;       // K: targets in Morton order (item order)
  const uint32_t *item_start;  // K + 1 by Morton rank
  const double *partial;
  const int *node_of_slot;
  double *rho, *j;
};
__global__ void k_combine_items(CombineArgs g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.K * g.NS) return;
  const int r = (int)(t / g.NS), slot = (int)(t % g.NS);
  const int node = g.node_of_slot[slot];
  if (node < 0) return;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t i = g.item_start[r]; i < g.item_start[r + 1]; ++i) {
    const double *p = g.partial + ((size_t)i * g.NS + slot) * 4;
    for (int k = 0; k < 4; ++k) v[k] += p[k];
  }
  const size_t gn = (size_t)g.order[r] * g.Np + node;
  if (g.rho) g.rho[gn] = v[0];
  if (g.j)
    for (int k = 0; k < g.D; ++k) g.j[gn * g.D + k] = v[1 + k];
}

// Work items of the list gather: target rank r (Morton order) gets
// ceil(n_live / 64) items of up to 64 list entries each.
__global__ void k_item_counts(int K, const uint32_t *__restrict__ order,
                              const uint32_t *__restrict__ n_live, uint32_t *__restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < K) cnt[r] = (n_live[order[r]] + 63u) / 64u;
}
__global__ void k_item_fill(int K, const uint32_t *__restrict__ order,
                            const uint32_t *__restrict__ item_start,
                            uint2 *__restrict__ items) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= K) return;
  const uint32_t e = order[r];
  for (uint32_t i = item_start[r], k = 0; i < item_start[r + 1]; ++i, ++k)
    items[i] = make_uint2(e, 64u * k);
}

// Sort key of every work item: the chunk of its first list entry, shifted
// right by `shift` (at most 16 key bits: two 8-bit sort passes; items whose
// first chunks share a key keep their target order).  The list gather runs the
// items in this order, so that the items reading a chunk (from ~200 targets)
// run at about the same time and share it in the L2.
__global__ void k_item_keys(uint32_t n_items, const uint2 *__restrict__ items,
                            const uint32_t *__restrict__ cap_off, const uint2 *__restrict__ ent,
                            int shift, uint32_t *__restrict__ key) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_items) return;
  const uint2 it = items[i];
  key[i] = ent[cap_off[it.x] + it.y].x >> shift;
}

// Targets that may receive charge: those whose candidate list holds a source
// element with particles, found through the transposed lists (one wave per
// source element; concurrent writes store the same value).
__global__ __launch_bounds__(256) void k_mark_active(int K, const uint32_t *__restrict__ ch_start,
                                                     const uint32_t *__restrict__ tr_start,
                                                     const uint32_t *__restrict__ tr,
                                                     uint8_t *__restrict__ active) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= K || ch_start[s + 1] == ch_start[s]) return;
  for (uint32_t k = tr_start[s] + lane; k < tr_start[s + 1]; k += 64) active[tr[k]] = 1;
}

// Candidate chunks of every target (the capacity of its chunk list): one wave
// per target sums the chunk counts of its candidate source elements.
__global__ __launch_bounds__(256) void k_cand_chunks(int K, const uint32_t *__restrict__ cand_start,
                                                     const uint32_t *__restrict__ cand,
                                                     const uint32_t *__restrict__ ch_start,
                                                     const uint8_t *__restrict__ active,
                                                     uint32_t *__restrict__ cap) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= K) return;
  if (!active[e]) {   // no candidate source element holds a particle
    if (lane == 0) cap[e] = 0;
    return;
  }
  uint32_t sum = 0;
  for (uint32_t k = cand_start[e] + lane; k < cand_start[e + 1]; k += 64) {
    const uint32_t s = cand[k];
    sum += ch_start[s + 1] - ch_start[s];
  }
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if (lane == 0) cap[e] = sum;
}
}  // namespace pyr

using namespace pyr;

// ---------------------------------------------------------------------------
// per-mesh deposition cache
// ---------------------------------------------------------------------------
struct DepositCache {
  double radius = -1;
  int64_t n_cand = 0;
  DevBuf cand_start, cand, cand_shift, geom, tab64, tab32, tab_R;
  double tab_radius = -1;
  // transposed lists (per source element) for the adjoint
  DevBuf tr_start, tr, tr_shift;
  double tr_radius = -1;
  // sort / staging scratch (sort_tmp: the library sort's temporary storage)
  DevBuf keys_a, keys_b, vals_a, vals_b, hist, hist_scan, sort_tmp, seg, tmp, pos, wgt, stats;
  // matrix-core gather over culled chunk lists: targets in Morton order,
  // grouped node tables, per-deposition chunk staging and the lists
  DevBuf t_order, tab32g, tab64g, node_slot, ch_meta, pos32c, wgt32c;
  DevBuf cap, ent, ent_off, n_live, items, partial;
  DevBuf pos64t, wgt64t;   // operands of the FP64 matrix-core gather
  DevBuf ikey, ik_a, ik_b, iv_a, iv_b;   // work items sorted by their first chunk
  DevBuf active;                         // targets that may receive charge
  int64_t max_indeg = 0;   // most targets whose candidate list holds one source
  // the lists above are those of the last deposition when list_valid
  bool list_valid = false;
  uint32_t list_items = 0;
  DevBuf list_stats;
  DevBuf nrm;            // normalized deposition: integrals, scaled charges, counter
  DevBuf skeys;   // Morton sort keys
  DevBuf chunks;         // per-particle gathers: 64-particle chunks (element, start),
                         // per-element chunk counts/offsets, total
  // the last sort's permutation (in vals_a or vals_b) and its particle count
  const uint32_t *last_perm = nullptr;
  int64_t last_perm_n = -1;
  std::vector<double> h_geom;   // K x 8
  // counts read back during a deposition (pinned), and the event after each
  // copy: the host waits on the event while later independent work runs
  uint32_t *h_cnt = nullptr;
  hipEvent_t ev_cnt[2] = {nullptr, nullptr};
};

// Reads a device count through pinned memory: the copy and an event go on the
// stream now; count_wait returns the value once the copy has landed, without
// waiting for work queued after it.
static void count_post(DepositCache *c, int k, const uint32_t *dev_count, hipStream_t s) {
  if (!c->h_cnt) PYR_HIP(hipHostMalloc((void **)&c->h_cnt, 4 * sizeof(uint32_t)));
  if (!c->ev_cnt[k]) PYR_HIP(hipEventCreateWithFlags(&c->ev_cnt[k], hipEventDisableTiming));
  PYR_HIP(hipMemcpyAsync(c->h_cnt + k, dev_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  PYR_HIP(hipEventRecord(c->ev_cnt[k], s));
}
static uint32_t count_wait(DepositCache *c, int k) {
  PYR_HIP(hipEventSynchronize(c->ev_cnt[k]));
  return c->h_cnt[k];
}

void pyr_deposit_cache_free(pyr_mesh *mesh) {
  if (!mesh || !mesh->dep) return;
  DepositCache *c = mesh->dep;
  for (DevBuf *b : {&c->cand_start, &c->cand, &c->cand_shift, &c->geom, &c->tab64, &c->tab32,
                    &c->tab_R, &c->tr_start, &c->tr, &c->tr_shift, &c->keys_a, &c->keys_b,
                    &c->vals_a, &c->vals_b, &c->skeys, &c->hist, &c->hist_scan,
                    &c->sort_tmp, &c->seg, &c->tmp, &c->pos, &c->wgt, &c->stats,
                    &c->nrm, &c->chunks, &c->t_order, &c->tab32g, &c->tab64g,
                    &c->node_slot, &c->ch_meta, &c->pos32c, &c->wgt32c, &c->cap,
                    &c->ent, &c->ent_off, &c->n_live, &c->items, &c->partial, &c->ikey,
                    &c->pos64t, &c->wgt64t,
                    &c->ik_a, &c->ik_b, &c->iv_a, &c->iv_b, &c->active, &c->list_stats})
    b->release();
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  for (hipEvent_t ev : c->ev_cnt)
    if (ev) (void)hipEventDestroy(ev);
  delete c;
  mesh->dep = nullptr;
}

// Host mirror of the mesh arrays the candidate builder needs.
struct HostGeom {
  int D, K, Np;
  std::vector<double> verts;      // K x (D+1) x D (element vertex coordinates)
  std::vector<double> fn, fr;     // face normals, rhs
  std::vector<uint32_t> nb, pax;  // face neighbours, periodic axes
  std::vector<double> nodes;      // K*Np x D
  double per_min[3], per_max[3];
};

static HostGeom download_geom(pyr_mesh *h) {
  HostGeom g;
  const int D = h->dim, K = h->K, F = D + 1;
  g.D = D;
  g.K = K;
  g.Np = h->Np;
  std::vector<uint32_t> elv((size_t)K * F);
  std::vector<double> vtx((size_t)h->nv * D);
  g.fn.resize((size_t)K * F * D);
  g.fr.resize((size_t)K * F);
  g.nb.resize((size_t)K * F);
  g.pax.resize((size_t)K * F);
  g.nodes.resize((size_t)K * h->Np * D);
  const DevMesh &m = h->dev;
  PYR_HIP(hipMemcpy(elv.data(), m.el_vertices, elv.size() * 4, hipMemcpyDeviceToHost));
  PYR_HIP(hipMemcpy(vtx.data(), m.vertices, vtx.size() * 8, hipMemcpyDeviceToHost));
  PYR_HIP(hipMemcpy(g.fn.data(), m.face_normal, g.fn.size() * 8, hipMemcpyDeviceToHost));
  PYR_HIP(hipMemcpy(g.fr.data(), m.face_rhs, g.fr.size() * 8, hipMemcpyDeviceToHost));
  PYR_HIP(hipMemcpy(g.nb.data(), m.face_neighbor, g.nb.size() * 4, hipMemcpyDeviceToHost));
  PYR_HIP(hipMemcpy(g.pax.data(), m.face_per_axis, g.pax.size() * 4, hipMemcpyDeviceToHost));
  PYR_HIP(hipMemcpy(g.nodes.data(), m.nodes, g.nodes.size() * 8, hipMemcpyDeviceToHost));
  g.verts.resize((size_t)K * F * D);
  for (int e = 0; e < K; ++e)
    for (int k = 0; k < F; ++k)
      for (int i = 0; i < D; ++i)
        g.verts[((size_t)e * F + k) * D + i] = vtx[(size_t)elv[(size_t)e * F + k] * D + i];
  for (int i = 0; i < 3; ++i) {
    g.per_min[i] = m.per_min[i];
    g.per_max[i] = m.per_max[i];
  }
  return g;
}

// Spatial grouping of an element's nodes into groups of four (the node groups
// of the matrix-core gather): in barycentric coordinates of element 0 mapped
// onto a regular simplex, repeatedly take the remaining node farthest from the
// centre together with its three nearest remaining nodes.  Every element
// holds the same unit nodes, so one grouping serves all.  Returns the node of
// each slot 4 g + i (-1: padding of the last group).
static std::vector<int> group_nodes(const HostGeom &g) {
  const int D = g.D, F = D + 1, np = g.Np;
  const int ng = (np + 3) / 4;
  // regular simplex vertices
  const double reg2[3][3] = {{0, 0, 0}, {1, 0, 0}, {0.5, 0.8660254037844386, 0}};
  const double reg3[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  const double *v0 = &g.verts[0];
  // barycentric coordinates from the edge matrix T (columns v_k - v_0)
  double T[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = 1; k < F; ++k)
    for (int i = 0; i < D; ++i) T[i][k - 1] = g.verts[(size_t)k * D + i] - v0[i];
  // inverse of T (D x D) by Gauss-Jordan
  double a[3][6];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < 2 * D; ++j) a[i][j] = j < D ? T[i][j] : (j - D == i ? 1.0 : 0.0);
  for (int c = 0; c < D; ++c) {
    int piv = c;
    for (int r = c + 1; r < D; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    for (int j = 0; j < 2 * D; ++j) std::swap(a[c][j], a[piv][j]);
    const double s = 1.0 / a[c][c];
    for (int j = 0; j < 2 * D; ++j) a[c][j] *= s;
    for (int r = 0; r < D; ++r)
      if (r != c) {
        const double f = a[r][c];
        for (int j = 0; j < 2 * D; ++j) a[r][j] -= f * a[c][j];
      }
  }
  std::vector<std::array<double, 3>> y(np);
  for (int n = 0; n < np; ++n) {
    double lam[4] = {1, 0, 0, 0};
    for (int i = 0; i < D; ++i) {
      double l = 0;
      for (int j = 0; j < D; ++j) l += a[i][D + j] * (g.nodes[(size_t)n * D + j] - v0[j]);
      lam[i + 1] = l;
      lam[0] -= l;
    }
    for (int k = 0; k < 3; ++k) {
      double s = 0;
      for (int v = 0; v < F; ++v) s += lam[v] * (D == 2 ? reg2[v][k] : reg3[v][k]);
      y[n][k] = s;
    }
  }
  double ctr[3] = {0, 0, 0};
  for (int n = 0; n < np; ++n)
    for (int k = 0; k < 3; ++k) ctr[k] += y[n][k] / np;
  auto dist2 = [&](const double *p, const double *q) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += (p[k] - q[k]) * (p[k] - q[k]);
    return s;
  };
  std::vector<int> slots(4 * (size_t)ng, -1);
  std::vector<char> used(np, 0);
  for (int grp = 0; grp < ng; ++grp) {
    int far = -1;
    double best = -1;
    for (int n = 0; n < np; ++n)
      if (!used[n] && dist2(y[n].data(), ctr) > best + 1e-12) {
        best = dist2(y[n].data(), ctr);
        far = n;
      }
    if (far < 0) break;
    used[far] = 1;
    slots[4 * grp] = far;
    for (int i = 1; i < 4; ++i) {
      int near = -1;
      double bd = 0;
      for (int n = 0; n < np; ++n)
        if (!used[n] && (near < 0 || dist2(y[n].data(), y[far].data()) < bd - 1e-12)) {
          near = n;
          bd = dist2(y[n].data(), y[far].data());
        }
      if (near < 0) break;
      used[near] = 1;
      slots[4 * grp + i] = near;
    }
  }
  return slots;
}

// Targets in Morton order of their centroids (10 bits per axis over the
// bounding box of the centroids): the list gather's processing order, so that
// the waves in flight on one XCD work on nearby targets and share chunks.
static std::vector<uint32_t> morton_order(const HostGeom &g, const std::vector<double> &cen) {
  const int D = g.D, K = g.K;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int e = 0; e < K; ++e)
    for (int i = 0; i < D; ++i) {
      lo[i] = std::min(lo[i], cen[(size_t)e * 3 + i]);
      hi[i] = std::max(hi[i], cen[(size_t)e * 3 + i]);
    }
  std::vector<std::pair<uint64_t, uint32_t>> key(K);
  for (int e = 0; e < K; ++e) {
    uint64_t code = 0;
    for (int i = 0; i < D; ++i) {
      const double span = hi[i] > lo[i] ? hi[i] - lo[i] : 1.0;
      const uint64_t q = (uint64_t)std::min(
          1023.0, std::max(0.0, (cen[(size_t)e * 3 + i] - lo[i]) / span * 1024.0));
      for (int b = 0; b < 10; ++b) code |= ((q >> b) & 1u) << (b * D + i);
    }
    key[e] = {code, (uint32_t)e};
  }
  std::sort(key.begin(), key.end());
  std::vector<uint32_t> order(K);
  for (int e = 0; e < K; ++e) order[e] = key[e].second;
  return order;
}

// Builds, for every target element, the list of (source element, periodic
// shift) pairs whose particles may lie within R of one of its nodes.  A BFS
// over face neighbours expands only accepted elements; the region within R of
// a convex element is convex, so every element meeting it is reached.
static void build_candidates(pyr_mesh *h, DepositCache *c, double R) {
  HostGeom g = download_geom(h);
  const int D = g.D, K = g.K, F = D + 1, np = g.Np;
  // per element: vertex centroid, vertex radius, node bound
  std::vector<double> cen((size_t)K * 3, 0.0), vrad(K, 0.0);
  c->h_geom.assign((size_t)K * 8, 0.0);
  for (int e = 0; e < K; ++e) {
    for (int k = 0; k < F; ++k)
      for (int i = 0; i < D; ++i) cen[(size_t)e * 3 + i] += g.verts[((size_t)e * F + k) * D + i] / F;
    double r = 0;
    for (int k = 0; k < F; ++k) {
      double s = 0;
      for (int i = 0; i < D; ++i) {
        double dv = g.verts[((size_t)e * F + k) * D + i] - cen[(size_t)e * 3 + i];
        s += dv * dv;
      }
      r = std::max(r, std::sqrt(s));
    }
    vrad[e] = r;
    double rn = 0;
    for (int i = 0; i < np; ++i) {
      double s = 0;
      for (int k = 0; k < D; ++k) {
        double dv = g.nodes[((size_t)e * np + i) * D + k] - cen[(size_t)e * 3 + k];
        s += dv * dv;
      }
      rn = std::max(rn, std::sqrt(s));
    }
    for (int i = 0; i < 3; ++i) c->h_geom[(size_t)e * 8 + i] = cen[(size_t)e * 3 + i];
    const double b = (R + rn) * (1 + 1e-12) + 1e-300;
    c->h_geom[(size_t)e * 8 + 3] = b * b;
    c->h_geom[(size_t)e * 8 + 4] = vrad[e];
  }
  const double slack = 1e-9 * R;
  auto accept = [&](int e, int s, const double *sh) {
    // bounding spheres
    double dd = 0;
    for (int i = 0; i < D; ++i) {
      double dv = cen[(size_t)s * 3 + i] + sh[i] - cen[(size_t)e * 3 + i];
      dd += dv * dv;
    }
    if (std::sqrt(dd) > R + vrad[e] + vrad[s] + slack) return false;
    // face planes of e against the vertices of s, and vice versa
    for (int f = 0; f < F; ++f) {
      const double *n = &g.fn[((size_t)e * F + f) * D];
      double mn = INFINITY;
      for (int k = 0; k < F; ++k) {
        double ip = 0;
        for (int i = 0; i < D; ++i) ip += n[i] * (g.verts[((size_t)s * F + k) * D + i] + sh[i]);
        mn = std::min(mn, ip);
      }
      if (mn - g.fr[(size_t)e * F + f] > R + slack) return false;
    }
    for (int f = 0; f < F; ++f) {
      const double *n = &g.fn[((size_t)s * F + f) * D];
      double rhs = g.fr[(size_t)s * F + f];
      for (int i = 0; i < D; ++i) rhs += n[i] * sh[i];
      double mn = INFINITY;
      for (int k = 0; k < F; ++k) {
        double ip = 0;
        for (int i = 0; i < D; ++i) ip += n[i] * g.verts[((size_t)e * F + k) * D + i];
        mn = std::min(mn, ip);
      }
      if (mn - rhs > R + slack) return false;
    }
    return true;
  };
  const int nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<uint32_t>> lists(K);
  std::vector<std::vector<uint8_t>> shifts(K);
  std::atomic<int> next{0};
  std::atomic<int> failed{0};
  auto worker = [&]() {
    std::vector<int> stamp(K, -1);
    std::vector<uint8_t> stamp_code(K, 0);
    std::vector<std::pair<uint32_t, uint8_t>> queue;
    for (;;) {
      int e = next.fetch_add(1);
      if (e >= K) break;
      queue.clear();
      queue.emplace_back((uint32_t)e, 0);
      stamp[e] = e;
      stamp_code[e] = 0;
      std::vector<std::pair<uint32_t, uint8_t>> acc;
      for (size_t qi = 0; qi < queue.size(); ++qi) {
        const uint32_t s = queue[qi].first;
        const uint8_t code = queue[qi].second;
        double sh[3] = {0, 0, 0};
        for (int a = 0; a < D; ++a) {
          int cc = (code >> (2 * a)) & 3;
          sh[a] = cc == 1 ? (g.per_max[a] - g.per_min[a])
                          : (cc == 2 ? -(g.per_max[a] - g.per_min[a]) : 0.0);
        }
        if (!accept(e, (int)s, sh)) continue;
        acc.emplace_back(s, code);
        for (int f = 0; f < F; ++f) {
          const uint32_t nb = g.nb[(size_t)s * F + f];
          if (nb == INVALID) continue;
          uint8_t ncode = code;
          const uint32_t ax = g.pax[(size_t)s * F + f];
          if (ax != INVALID) {
            // crossing a periodic face: the neighbour sits one period away
            const double L = g.per_max[ax] - g.per_min[ax];
            double fcx = 0;
            for (int k = 0; k < F; ++k) fcx += g.verts[((size_t)s * F + k) * D + ax] / F;
            const bool at_max = fcx > 0.5 * (g.per_min[ax] + g.per_max[ax]);
            int cur = (code >> (2 * ax)) & 3;
            int val = cur == 1 ? 1 : (cur == 2 ? -1 : 0);
            val += at_max ? 1 : -1;
            if (val < -1 || val > 1) {
              failed = 1;
              continue;
            }
            (void)L;
            ncode = (uint8_t)((code & ~(3u << (2 * ax))) |
                              ((val == 1 ? 1u : (val == -1 ? 2u : 0u)) << (2 * ax)));
          }
          if (stamp[nb] == e) {
            if (stamp_code[nb] != ncode) failed = 1;
            continue;
          }
          stamp[nb] = e;
          stamp_code[nb] = ncode;
          queue.emplace_back(nb, ncode);
        }
      }
      std::sort(acc.begin(), acc.end());
      lists[e].reserve(acc.size());
      shifts[e].reserve(acc.size());
      for (auto &p : acc) {
        lists[e].push_back(p.first);
        shifts[e].push_back(p.second);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; ++t) pool.emplace_back(worker);
  for (auto &t : pool) t.join();
  if (failed)
    throw Error(PYR_ERR_INVALID_ARG,
                "shape radius reaches the same element through two periodic images; the "
                "period must exceed the shape support");
  std::vector<uint32_t> starts((size_t)K + 1, 0);
  for (int e = 0; e < K; ++e) starts[e + 1] = starts[e] + (uint32_t)lists[e].size();
  std::vector<uint32_t> flat(starts[K]);
  std::vector<uint8_t> flat_s(starts[K]);
  for (int e = 0; e < K; ++e) {
    std::copy(lists[e].begin(), lists[e].end(), flat.begin() + starts[e]);
    std::copy(shifts[e].begin(), shifts[e].end(), flat_s.begin() + starts[e]);
  }
  upload(c->cand_start, starts.data(), starts.size());
  upload(c->cand, flat.data(), std::max<size_t>(1, flat.size()));
  upload(c->cand_shift, flat_s.data(), std::max<size_t>(1, flat_s.size()));
  upload(c->geom, c->h_geom.data(), c->h_geom.size());
  c->n_cand = starts[K];
  const std::vector<uint32_t> order = morton_order(g, cen);
  upload(c->t_order, order.data(), order.size());
  std::vector<uint32_t> indeg(K, 0);
  for (uint32_t s : flat) indeg[s]++;
  c->max_indeg = K ? *std::max_element(indeg.begin(), indeg.end()) : 0;
  const std::vector<int> slots = group_nodes(g);
  upload(c->node_slot, slots.data(), slots.size());
  c->radius = R;
}

static unsigned blocks_for(int64_t n, int bs = 256) {
  return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs);
}

// Stable radix sort of the indices 0..n-1 by keys clamped to [0, kmax];
// returns the sorted indices (in va or vb) and, with sorted_keys, the sorted
// keys.  Large sorts (the particles) run on rocPRIM's onesweep radix sort;
// small ones (the work items) as 8-bit LSD passes of k_rs_hist / k_rs_scatter,
// a few launches each where rocPRIM would take a block sort and a chain of
// merge passes.  Both are stable, so the permutation is the same.
constexpr int64_t SORT_LIBRARY_MIN = 1 << 20;

static const uint32_t *radix_sort(DepositCache *c, int64_t n, const uint32_t *keys,
                                  uint32_t kmax, DevBuf &ka, DevBuf &kb, DevBuf &va,
                                  DevBuf &vb, const uint32_t **sorted_keys = nullptr) {
  hipStream_t s = stream();
  if (n >= SORT_LIBRARY_MIN) {
    for (DevBuf *b : {&ka, &va}) b->reserve(sizeof(uint32_t) * (size_t)n);
    const uint32_t *perm =
        sort_pairs(c->sort_tmp, n, keys, kmax, ka.as<uint32_t>(), va.as<uint32_t>(), s);
    if (sorted_keys) *sorted_keys = ka.as<uint32_t>();
    return perm;
  }
  const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)kmax) ++bits;
  const int passes = (bits + 7) / 8;
  for (DevBuf *b : {&ka, &kb, &va, &vb}) b->reserve(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, n));
  c->hist.reserve(sizeof(uint32_t) * 256 * ntiles);
  c->hist_scan.reserve(sizeof(uint32_t) * 256 * ntiles);
  c->tmp.reserve(sizeof(uint32_t) * scan_tmp_words(256 * ntiles));
  const uint32_t *kin = keys, *vin = nullptr;
  uint32_t *kout = ka.as<uint32_t>(), *vout = va.as<uint32_t>();
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_BS), 0, s, kin, n, 8 * p,
                       c->hist.as<uint32_t>(), (int)ntiles, kmax);
    PYR_CHECK_LAUNCH();
    scan_exclusive_u32(c->hist.as<uint32_t>(), c->hist_scan.as<uint32_t>(), 256 * ntiles,
                       nullptr, c->tmp.as<uint32_t>(), s);
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(RS_BS), 0, s, kin, vin, kout,
                       vout, n, 8 * p, c->hist_scan.as<uint32_t>(), (int)ntiles, kmax);
    PYR_CHECK_LAUNCH();
    kin = kout;
    vin = vout;
    if (kout == ka.as<uint32_t>()) {
      kout = kb.as<uint32_t>();
      vout = vb.as<uint32_t>();
    } else {
      kout = ka.as<uint32_t>();
      vout = va.as<uint32_t>();
    }
  }
  if (sorted_keys) *sorted_keys = kin;
  return vin;
}

// Stable radix sort of particle indices by keys (element ids, or element ids
// followed by a Morton code when key_shift > 0); returns the permutation (in
// vals_a or vals_b) and leaves per-element segment starts in c->seg.
static const uint32_t *sort_by_element(pyr_mesh *h, DepositCache *c, int64_t n,
                                       const uint32_t *elem, const uint32_t *keys = nullptr,
                                       int key_shift = 0) {
  const int K = h->K;
  hipStream_t s = stream();
  c->seg.reserve(sizeof(uint32_t) * ((size_t)K + 2));
  if (!keys) {
    keys = elem;
    key_shift = 0;
  }
  const uint32_t *sorted = nullptr;
  c->last_perm_n = -1;
  const uint32_t *perm = radix_sort(c, n, keys, (uint32_t)K << key_shift, c->keys_a, c->keys_b,
                                    c->vals_a, c->vals_b, &sorted);
  hipLaunchKernelGGL(k_seg_from_sorted, dim3(blocks_for((int64_t)K + 1)), dim3(256), 0, s, K, n,
                     sorted, key_shift, c->seg.as<uint32_t>());
  PYR_CHECK_LAUNCH();
  c->last_perm = perm;
  c->last_perm_n = n;
  return perm;
}

namespace pyr {
const uint32_t *element_order(pyr_mesh *mesh, int64_t n, const uint32_t *elem) {
  if (!mesh->dep) mesh->dep = new DepositCache();
  return sort_by_element(mesh, mesh->dep, n, elem);
}

const uint32_t *deposit_last_order(pyr_mesh *mesh, int64_t n) {
  if (!mesh || !mesh->dep || n <= 0 || mesh->dep->last_perm_n != n) return nullptr;
  return mesh->dep->last_perm;
}
}  // namespace pyr

// Sorts the particles by (element, Morton code inside the element) and, with
// stage64, stages them in that order for the FP64 kernels; returns the
// permutation.  The Morton sub-order makes batches of consecutive particles
// spatially coherent (the FP64 gather's queued batches, the FP32C chunks,
// whose bounding spheres it keeps small).
// sorted: a permutation (and c->seg) left by an earlier sort of the same
// positions and elements, which is then only restaged.
static const uint32_t *sort_and_stage(pyr_mesh *mesh, DepositCache *c, int64_t n,
                                      const double *x, const double *q, const double *v,
                                      const uint32_t *elem, double normalizer,
                                      bool stage64 = true,
                                      const uint32_t *sorted = nullptr) {
  const int K = mesh->K, D = mesh->dim;
  hipStream_t s = stream();
  const uint32_t *perm = sorted;
  if (!perm) {
    const uint32_t *keys = nullptr;
    int shift = 0;
    if ((uint64_t)K < ((uint64_t)1 << (32 - MORTON_BITS))) {
      c->skeys.reserve(sizeof(uint32_t) * (size_t)n);
      if (D == 2)
        hipLaunchKernelGGL(k_sort_keys<2>, dim3(blocks_for(n)), dim3(256), 0, s, n, K, elem, x,
                           c->geom.as<double>(), c->skeys.as<uint32_t>());
      else
        hipLaunchKernelGGL(k_sort_keys<3>, dim3(blocks_for(n)), dim3(256), 0, s, n, K, elem, x,
                           c->geom.as<double>(), c->skeys.as<uint32_t>());
      PYR_CHECK_LAUNCH();
      keys = c->skeys.as<uint32_t>();
      shift = MORTON_BITS;
    }
    perm = sort_by_element(mesh, c, n, elem, keys, shift);
  }
  if (stage64) {
    c->pos.reserve(sizeof(double4) * (size_t)n);
    c->wgt.reserve(sizeof(double4) * (size_t)n);
    hipLaunchKernelGGL(k_stage, dim3(blocks_for(n)), dim3(256), 0, s, D, n, perm, x, q, v,
                       normalizer, c->pos.as<double4>(), c->wgt.as<double4>());
    PYR_CHECK_LAUNCH();
  }
  return perm;
}

// P = 2^ceil(log2 R): the FP32 gathers evaluate t / P <= 1 (exact scaling)
static double p_scale(double radius) {
  int p_exp = 0;
  const double p_mant = std::frexp(radius, &p_exp);
  return std::ldexp(1.0, p_mant == 0.5 ? p_exp - 1 : p_exp);
}

// candidate lists and node tables for radius R (cached per radius)
static void prepare_tables(pyr_mesh *mesh, DepositCache *c, double radius) {
  const int K = mesh->K, Np = mesh->Np, D = mesh->dim;
  hipStream_t s = stream();
  if (c->radius != radius) build_candidates(mesh, c, radius);
  if (c->tab_radius != radius) {
    c->tab64.reserve(sizeof(double) * (size_t)K * Np * 4);
    c->tab32.reserve(sizeof(float) * (size_t)K * Np * 4);
    hipLaunchKernelGGL(k_node_table<double>, dim3(blocks_for((int64_t)K * Np)), dim3(256), 0, s,
                       D, K, Np, mesh->dev.nodes, c->geom.as<double>(), 1.0 / radius,
                       c->tab64.as<double>());
    hipLaunchKernelGGL(k_node_table<float>, dim3(blocks_for((int64_t)K * Np)), dim3(256), 0, s,
                       D, K, Np, mesh->dev.nodes, c->geom.as<double>(), 1.0 / radius,
                       c->tab32.as<float>());
    PYR_CHECK_LAUNCH();
    const int ng = (Np + 3) / 4;
    c->tab32g.reserve(sizeof(float4) * (size_t)K * 4 * ng);
    c->tab64g.reserve(sizeof(double4) * (size_t)K * 4 * ng);
    hipLaunchKernelGGL(k_node_table_g, dim3(blocks_for((int64_t)K * ng)), dim3(256), 0, s, D, K,
                       Np, ng, mesh->dev.nodes, c->geom.as<double>(), 1.0 / radius,
                       1.0 / p_scale(radius), c->node_slot.as<int>(), c->tab32g.as<float4>(),
                       c->tab64g.as<double4>());
    PYR_CHECK_LAUNCH();
    c->tab_radius = radius;
  }
}

// Transpose of the candidate lists: for each source element, the target
// elements whose list holds it, in ascending target order, with the same
// shift code (the shift moves the source particle next to the target).
static void build_transpose(pyr_mesh *mesh, DepositCache *c) {
  const int K = mesh->K;
  const size_t nc = (size_t)c->n_cand;
  std::vector<uint32_t> start((size_t)K + 1), cand(std::max<size_t>(1, nc));
  std::vector<uint8_t> code(std::max<size_t>(1, nc));
  PYR_HIP(hipMemcpy(start.data(), c->cand_start.ptr, start.size() * 4, hipMemcpyDeviceToHost));
  if (nc) {
    PYR_HIP(hipMemcpy(cand.data(), c->cand.ptr, nc * 4, hipMemcpyDeviceToHost));
    PYR_HIP(hipMemcpy(code.data(), c->cand_shift.ptr, nc, hipMemcpyDeviceToHost));
  }
  std::vector<uint32_t> tstart((size_t)K + 1, 0);
  for (size_t k = 0; k < nc; ++k) tstart[cand[k] + 1]++;
  for (int e = 0; e < K; ++e) tstart[e + 1] += tstart[e];
  std::vector<uint32_t> fill(tstart.begin(), tstart.end() - 1);
  std::vector<uint32_t> tr(std::max<size_t>(1, nc));
  std::vector<uint8_t> tcode(std::max<size_t>(1, nc));
  for (int e = 0; e < K; ++e)
    for (uint32_t k = start[e]; k < start[e + 1]; ++k) {
      const uint32_t src = cand[k];
      const uint32_t at = fill[src]++;
      tr[at] = (uint32_t)e;
      tcode[at] = code[k];
    }
  upload(c->tr_start, tstart.data(), tstart.size());
  upload(c->tr, tr.data(), tr.size());
  upload(c->tr_shift, tcode.data(), tcode.size());
  c->tr_radius = c->radius;
}


// Work list of the per-particle gathers (shape integral, averaging, adjoint):
// the particles of each source element in chunks of up to 64 (one wave each),
// so that a dense element spreads over many waves.
__global__ void k_chunk_counts(int K, const uint32_t *__restrict__ seg,
                               uint32_t *__restrict__ nch) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < K) nch[s] = (seg[s + 1] - seg[s] + 63u) / 64u;
}

__global__ void k_chunk_fill(int K, const uint32_t *__restrict__ seg,
                             const uint32_t *__restrict__ nch, const uint32_t *__restrict__ off,
                             uint2 *__restrict__ chunks) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= K) return;
  const uint32_t o = off[s], p0 = seg[s];
  for (uint32_t k = 0; k < nch[s]; ++k) chunks[o + k] = make_uint2((uint32_t)s, p0 + 64u * k);
}

// Builds the chunk list of the current sort (c->seg); returns the launch grid
// (a multiple of the XCD grouping, an upper bound on the chunk count).
static unsigned build_chunks(pyr_mesh *mesh, DepositCache *c, int64_t n, ChunkList &cl) {
  const int K = mesh->K;
  hipStream_t s = stream();
  const int64_t bound = (n + 63) / 64 + std::min<int64_t>(K, n);
  const size_t words = (size_t)2 * bound + 2 * (size_t)K + 1;
  c->chunks.reserve(sizeof(uint32_t) * words);
  uint2 *chunks = c->chunks.as<uint2>();
  uint32_t *nch = reinterpret_cast<uint32_t *>(chunks + bound);
  uint32_t *off = nch + K;
  uint32_t *total = off + K;
  c->tmp.reserve(sizeof(uint32_t) * scan_tmp_words(K));
  hipLaunchKernelGGL(k_chunk_counts, dim3(blocks_for(K)), dim3(256), 0, s, K,
                     c->seg.as<uint32_t>(), nch);
  PYR_CHECK_LAUNCH();
  scan_exclusive_u32(nch, off, K, total, c->tmp.as<uint32_t>(), s);
  hipLaunchKernelGGL(k_chunk_fill, dim3(blocks_for(K)), dim3(256), 0, s, K,
                     c->seg.as<uint32_t>(), nch, off, chunks);
  PYR_CHECK_LAUNCH();
  cl.chunks = chunks;
  cl.n_chunks = total;
  return gather_grid((int)bound);
}

// q / I per particle for the normalized deposition; located particles whose
// shape integral vanishes deposit nothing and are counted (the reference warns
// for each, dep_normshape.hpp:175-183).  scale (optional) receives q / I.
__global__ void k_normalized_charges(int64_t n, int K, const uint32_t *__restrict__ elem,
                                     const double *__restrict__ q,
                                     const double *__restrict__ integral,
                                     double *__restrict__ qs, double *scale,
                                     unsigned long long *zeros) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool located = elem[i] < (uint32_t)K;
  const double I = integral[i];
  const double v = (located && I != 0.0) ? q[i] / I : 0.0;
  qs[i] = v;
  if (scale) scale[i] = v;
  if (located && I == 0.0) atomicAdd(zeros, 1ull);
}

static void deposit_impl(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                         int precision, int64_t n, int vdim, const double *x, const double *q,
                         const double *v, const uint32_t *elem, double *rho, double *j,
                         const uint32_t *sorted = nullptr) {
  {
    PYR_REQUIRE(mesh, "null mesh");
    PYR_REQUIRE(radius > 0 && std::isfinite(radius), "shape radius must be positive");
    PYR_REQUIRE(alpha > 0, "shape exponent must be positive");
    PYR_REQUIRE(precision == PYR_FP64 || precision == PYR_FP32C, "unknown precision");
    PYR_REQUIRE(vdim == mesh->dim, "vdim must equal the mesh dimension");
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    PYR_REQUIRE(!j || v || n == 0, "j requires velocities");
    if (!rho && !j) return;
    if (!mesh->dep) mesh->dep = new DepositCache();
    DepositCache *c = mesh->dep;
    const int K = mesh->K, Np = mesh->Np, D = mesh->dim;
    hipStream_t s = stream();
    prepare_tables(mesh, c, radius);
    const uint32_t *perm = nullptr;
    if (n > 0) PYR_REQUIRE(x && q && elem, "null pointer");
    {
      ProfScope prof(PROF_SORT);
      if (n > 0) {
        // the FP64 staging of the sorted order serves the pair-counting mode's
        // candidate gather; the list gathers stage their own chunks
        const bool stage64 = precision == PYR_FP64 && mesh->deposit_count != 0;
        perm = sort_and_stage(mesh, c, n, x, q, j ? v : nullptr, elem, normalizer, stage64,
                              sorted);
      } else {
        c->seg.reserve(sizeof(uint32_t) * ((size_t)K + 1));
        PYR_HIP(hipMemsetAsync(c->seg.ptr, 0, sizeof(uint32_t) * ((size_t)K + 1), s));
      }
    }
    GatherArgs ga{};
    ga.K = K;
    ga.Np = Np;
    ga.R = radius;
    ga.invR = 1.0 / radius;
    ga.R2 = radius * radius;
    ga.alpha = alpha;
    for (int a = 0; a < 3; ++a)
      ga.shift_len[a] = mesh->dev.per_max[a] - mesh->dev.per_min[a];
    ga.cand_start = c->cand_start.as<uint32_t>();
    ga.cand = c->cand.as<uint32_t>();
    ga.cand_shift = c->cand_shift.as<uint8_t>();
    ga.seg = c->seg.as<uint32_t>();
    ga.pos = c->pos.as<double4>();
    ga.wgt = c->wgt.as<double4>();
    ga.geom = c->geom.as<double>();
    ga.face_normal = mesh->dev.face_normal;
    ga.face_rhs = mesh->dev.face_rhs;
    ga.node_tab = precision == PYR_FP64 ? (const void *)c->tab64.ptr : (const void *)c->tab32.ptr;
    ga.rho = rho;
    ga.j = j;
    c->stats.reserve(4 * sizeof(unsigned long long));
    ga.stats = c->stats.as<unsigned long long>();
    const bool count = mesh->deposit_count != 0;
    if (count) PYR_HIP(hipMemsetAsync(ga.stats, 0, 4 * sizeof(unsigned long long), s));
    if (precision == PYR_FP32C || !count) {
      if (count) throw Error(PYR_ERR_INVALID_ARG, "deposit counting needs the FP64 precision");
      // chunks of up to 64 sorted particles per source element, staged in
      // FP32 with their bounding spheres, per-target culled chunk lists, then
      // the list gather (FP32C: matrix cores; FP64: the FP64 staging)
      ChunkList cl{};
      const uint32_t *ch_start = nullptr;
      {
        ProfScope prof(PROF_SORT);
        build_chunks(mesh, c, n, cl);
        // build_chunks lays out chunks[bound], counts[K], offsets[K], total:
        // the offsets followed by the total are the K + 1 starts
        ch_start = cl.n_chunks - K;
        count_post(c, 0, cl.n_chunks, s);
      }
      // one scratch: capacities (later item counts) K, list offsets K + 1, scan
      // partials, item starts K + 1
      c->cap.reserve(sizeof(uint32_t) * ((size_t)3 * K + 3 + scan_tmp_words(K)));
      uint32_t *cap = c->cap.as<uint32_t>(), *cap_off = cap + K + 1;
      uint32_t *scan_tmp = cap_off + K + 1;
      uint32_t *istart = scan_tmp + scan_tmp_words(K);
      {
        // the list capacities need no chunk count: queued before the host
        // waits for it
        ProfScope prof(PROF_SORT);
        if (c->tr_radius != radius) build_transpose(mesh, c);
        c->active.reserve((size_t)std::max(1, K));
        PYR_HIP(hipMemsetAsync(c->active.ptr, 0, (size_t)K, s));
        hipLaunchKernelGGL(k_mark_active, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, s, K,
                           ch_start, c->tr_start.as<uint32_t>(), c->tr.as<uint32_t>(),
                           c->active.as<uint8_t>());
        PYR_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_cand_chunks, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, s, K,
                           c->cand_start.as<uint32_t>(), c->cand.as<uint32_t>(), ch_start,
                           c->active.as<uint8_t>(), cap);
        PYR_CHECK_LAUNCH();
        scan_exclusive_u32(cap, cap_off, K, cap_off + K, scan_tmp, s);
      }
      int64_t nch = 0;
      {
        ProfScope prof(PROF_SORT);
        nch = count_wait(c, 0);
        c->ch_meta.reserve(sizeof(float4) * (size_t)std::max<int64_t>(1, nch));
        if (precision == PYR_FP32C) {
          // chunk nch: all zero, read by the list gather past the end of a list
          c->pos32c.reserve(sizeof(float4) * 64 * (size_t)(nch + 1));
          c->wgt32c.reserve(sizeof(float4) * 64 * (size_t)(nch + 1));
          PYR_HIP(hipMemsetAsync(c->pos32c.as<float4>() + 64 * (size_t)nch, 0,
                                 sizeof(float4) * 64, s));
          PYR_HIP(hipMemsetAsync(c->wgt32c.as<float4>() + 64 * (size_t)nch, 0,
                                 sizeof(float4) * 64, s));
          if (nch > 0) {
            const double scale = normalizer * std::pow(p_scale(radius), alpha);
            hipLaunchKernelGGL(k_stage32c, dim3((unsigned)((nch + 3) / 4)), dim3(256), 0, s, D,
                               nch, cl.chunks, c->seg.as<uint32_t>(), perm, x, q,
                               j ? v : nullptr, scale, c->geom.as<double>(),
                               c->pos32c.as<float4>(), c->wgt32c.as<float>(),
                               c->ch_meta.as<float4>());
            PYR_CHECK_LAUNCH();
          }
        } else {
          c->pos64t.reserve(sizeof(double) * 256 * (size_t)std::max<int64_t>(1, nch));
          c->wgt64t.reserve(sizeof(double) * 256 * (size_t)std::max<int64_t>(1, nch));
          if (nch > 0) {
            hipLaunchKernelGGL(k_stage64c, dim3((unsigned)((nch + 3) / 4)), dim3(256), 0, s, D,
                               nch, cl.chunks, c->seg.as<uint32_t>(), perm, x, q,
                               j ? v : nullptr, normalizer, c->geom.as<double>(), 1.0 / radius,
                               c->pos64t.as<double>(), c->wgt64t.as<double>(),
                               c->ch_meta.as<float4>());
            PYR_CHECK_LAUNCH();
          }
        }
      }
      // per-target lists of the chunks that reach a node group: capacity =
      // candidate chunks (at most nch x max_indeg in total), then the cull
      const int64_t bound = std::max<int64_t>(1, nch * std::max<int64_t>(1, c->max_indeg));
      PYR_REQUIRE(bound < ((int64_t)1 << 32), "deposition chunk lists exceed 2^32 entries");
      c->ent.reserve(sizeof(uint2) * (size_t)bound);
      c->ent_off.reserve(sizeof(float4) * (size_t)bound);
      c->n_live.reserve(sizeof(uint32_t) * (size_t)std::max(1, K));
      {
        ProfScope prof(PROF_SORT);
        CullArgs cu{};
        cu.K = K;
        cu.R = radius;
        for (int a = 0; a < 3; ++a) cu.shift_len[a] = ga.shift_len[a];
        cu.cand_start = c->cand_start.as<uint32_t>();
        cu.cand = c->cand.as<uint32_t>();
        cu.cand_shift = c->cand_shift.as<uint8_t>();
        cu.ch_start = ch_start;
        cu.ch_meta = c->ch_meta.as<float4>();
        cu.geom = c->geom.as<double>();
        cu.ntab = c->tab32g.as<float4>();
        cu.node_slot = c->node_slot.as<int>();
        cu.inv_rp = (float)(1.0 / (radius * p_scale(radius)));
        cu.cap_off = cap_off;
        cu.ent = c->ent.as<uint2>();
        cu.ent_off = c->ent_off.as<float4>();
        cu.n_live = c->n_live.as<uint32_t>();
        cu.active = c->active.as<uint8_t>();
        launch_cull_chunks(D, Np, cu);
      }
      // work items of up to 64 list entries, filled in Morton order of the
      // targets, then ordered by their first chunk
      const int ns = 4 * ((Np + 3) / 4);
      uint32_t *icnt = cap;   // the capacities are no longer needed
      uint32_t n_items = 0;
      const uint32_t *item_order = nullptr;
      {
        ProfScope prof(PROF_SORT);
        hipLaunchKernelGGL(k_item_counts, dim3(blocks_for(K)), dim3(256), 0, s, K,
                           c->t_order.as<uint32_t>(), c->n_live.as<uint32_t>(), icnt);
        PYR_CHECK_LAUNCH();
        scan_exclusive_u32(icnt, istart, K, istart + K, scan_tmp, s);
        count_post(c, 1, istart + K, s);
        // the items fit in (list entries / 64 + targets): filled before the
        // host waits for their count
        const int64_t item_bound = (bound + 63) / 64 + K;
        c->items.reserve(sizeof(uint2) * (size_t)item_bound);
        hipLaunchKernelGGL(k_item_fill, dim3(blocks_for(K)), dim3(256), 0, s, K,
                           c->t_order.as<uint32_t>(), istart, c->items.as<uint2>());
        PYR_CHECK_LAUNCH();
        n_items = count_wait(c, 1);
        c->partial.reserve(sizeof(double) * 4 * ns * (size_t)std::max<uint32_t>(1, n_items));
        if (n_items) {
          c->ikey.reserve(sizeof(uint32_t) * (size_t)n_items);
          int ishift = 0;
          while (((uint64_t)std::max<int64_t>(1, nch) >> ishift) >= ((uint64_t)1 << 16)) ++ishift;
          hipLaunchKernelGGL(k_item_keys, dim3(blocks_for(n_items)), dim3(256), 0, s, n_items,
                             c->items.as<uint2>(), cap_off, c->ent.as<uint2>(), ishift,
                             c->ikey.as<uint32_t>());
          PYR_CHECK_LAUNCH();
          item_order = radix_sort(c, n_items, c->ikey.as<uint32_t>(),
                                  (uint32_t)(std::max<int64_t>(1, nch) >> ishift), c->ik_a,
                                  c->ik_b, c->iv_a, c->iv_b);
        }
      }
      ListArgs la{};
      la.K = K;
      la.R = radius;
      la.invR = 1.0 / radius;
      la.inv_p = (float)(1.0 / p_scale(radius));
      la.alpha = (float)alpha;
      la.items = c->items.as<uint2>();
      la.item_order = item_order;
      la.n_items = istart + K;
      la.cap_off = cap_off;
      la.n_live = c->n_live.as<uint32_t>();
      la.ent = c->ent.as<uint2>();
      la.ent_off = c->ent_off.as<float4>();
      la.pos = c->pos32c.as<float4>();
      la.wgt = c->wgt32c.as<float4>();
      la.ntab = c->tab32g.as<float4>();
      la.partial = c->partial.as<double>();
      la.zero_chunk = (uint32_t)nch;
      CombineArgs cb{};
      cb.K = K;
      cb.NS = ns;
      cb.Np = Np;
      cb.D = D;
      cb.order = c->t_order.as<uint32_t>();
      cb.item_start = istart;
      cb.partial = c->partial.as<double>();
      cb.node_of_slot = c->node_slot.as<int>();
      cb.rho = rho;
      cb.j = j;
      {
        ProfScope prof(PROF_GATHER);   // the list gather alone
        if (n_items && precision == PYR_FP32C) {
          launch_gather_cl(D, Np, la, n_items, alpha == 2.0);
        } else if (n_items) {
          ListArgsM64 lm{};
          lm.K = K;
          lm.R = radius;
          lm.invR = 1.0 / radius;
          lm.alpha = alpha;
          for (int a = 0; a < 3; ++a) lm.shift_len[a] = ga.shift_len[a];
          lm.items = la.items;
          lm.item_order = la.item_order;
          lm.n_items = la.n_items;
          lm.cap_off = cap_off;
          lm.n_live = la.n_live;
          lm.ent = la.ent;
          lm.chunks = cl.chunks;
          lm.geom = c->geom.as<double>();
          lm.node_tab = c->tab64g.as<double4>();
          lm.pos_t = c->pos64t.as<double>();
          lm.wgt_t = c->wgt64t.as<double>();
          lm.partial = c->partial.as<double>();
          launch_gather_mf64(D, Np, lm, n_items, shape_unit_square(alpha, radius));
        }
      }
      ProfScope prof(PROF_SORT);       // with the other per-deposition set-up work
      hipLaunchKernelGGL(k_combine_items, dim3(blocks_for((int64_t)K * ns)), dim3(256), 0, s,
                         cb);
      PYR_CHECK_LAUNCH();
      c->list_valid = precision == PYR_FP32C;
      c->list_items = n_items;
      return;
    }
    c->list_valid = false;
    ProfScope prof(PROF_GATHER);
    launch_gather_fp64(D, Np, ga, K, shape_unit_square(alpha, radius), count);
  }
}


// The averaging pusher's forces (push_average.hpp:338-443): one sort of the
// particles by element, then the shape-weighted gather k_shape_average.
// Stream-ordered, no host synchronization; returns the device counter of
// located particles with a zero shape integral (null when n = 0).  vis: the
// six vis arrays (pt_e, pt_b, pt_e_stddev, pt_b_stddev, el_force, mag_force)
// or null.
namespace pyr {
unsigned long long *average_forces_impl(pyr_mesh *mesh, double radius, double alpha,
                                        double normalizer, int64_t n, int vdim,
                                        const double *const fields[6], const double *x,
                                        const uint32_t *elem, const double *q, const double *v,
                                        const double *integral_weights, double *force,
                                        double *integral_out, double *const *vis) {
  PYR_REQUIRE(radius > 0 && std::isfinite(radius), "shape radius must be positive");
  PYR_REQUIRE(alpha > 0, "shape exponent must be positive");
  PYR_REQUIRE(vdim == mesh->dim, "vdim must equal the mesh dimension");
  PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
  PYR_REQUIRE(mesh->dev.jacobian, "the mesh was created without element jacobians");
  if (vis)
    for (int k = 0; k < 6; ++k) PYR_REQUIRE(vis[k], "all six vis arrays or none");
  if (n == 0) return nullptr;
  PYR_REQUIRE(x && q && v && elem && integral_weights && force, "null pointer");
  if (!mesh->dep) mesh->dep = new DepositCache();
  DepositCache *c = mesh->dep;
  const int K = mesh->K, Np = mesh->Np, D = mesh->dim;
  hipStream_t s = stream();
  prepare_tables(mesh, c, radius);
  if (c->tr_radius != radius) build_transpose(mesh, c);
  const uint32_t *perm;
  {
    ProfScope prof(PROF_SORT);
    perm = sort_and_stage(mesh, c, n, x, q, nullptr, elem, normalizer);
  }
  // particles without an element keep zero outputs
  PYR_HIP(hipMemsetAsync(force, 0, sizeof(double) * (size_t)n * vdim, s));
  if (integral_out) PYR_HIP(hipMemsetAsync(integral_out, 0, sizeof(double) * (size_t)n, s));
  if (vis) {
    const size_t w[6] = {3, 3, 1, 1, (size_t)vdim, (size_t)vdim};
    for (int k = 0; k < 6; ++k)
      PYR_HIP(hipMemsetAsync(vis[k], 0, sizeof(double) * (size_t)n * w[k], s));
  }
  c->nrm.reserve(sizeof(double) * ((size_t)2 * n + 1) + sizeof(unsigned long long));
  unsigned long long *zeros = c->nrm.as<unsigned long long>();
  PYR_HIP(hipMemsetAsync(zeros, 0, sizeof(unsigned long long), s));
  AverageArgs aa{};
  IntegralArgs &ia = aa.b;
  ia.K = K;
  ia.R = radius;
  ia.invR = 1.0 / radius;
  ia.alpha = alpha;
  ia.normalizer = normalizer;
  for (int a = 0; a < 3; ++a) ia.shift_len[a] = mesh->dev.per_max[a] - mesh->dev.per_min[a];
  ia.tr_start = c->tr_start.as<uint32_t>();
  ia.tr = c->tr.as<uint32_t>();
  ia.tr_shift = c->tr_shift.as<uint8_t>();
  ia.seg = c->seg.as<uint32_t>();
  ia.perm = perm;
  ia.pos = c->pos.as<double4>();
  ia.geom = c->geom.as<double>();
  ia.node_tab = c->tab64.as<double>();
  ia.iw = integral_weights;
  ia.jac = mesh->dev.jacobian;
  ia.integral = integral_out;
  for (int f = 0; f < 6; ++f) aa.f[f] = fields[f];
  aa.q = q;
  aa.v = v;
  aa.vdim = vdim;
  aa.force = force;
  if (vis) {
    aa.vis_e = vis[0];
    aa.vis_b = vis[1];
    aa.vis_e_std = vis[2];
    aa.vis_b_std = vis[3];
    aa.vis_el = vis[4];
    aa.vis_mag = vis[5];
  }
  aa.zeros = zeros;
  const unsigned grid = build_chunks(mesh, c, n, ia.cl);
  ProfScope prof(PROF_AVERAGE);
  launch_shape_average(D, Np, aa, grid, shape_unit_square(alpha, radius), vis != nullptr);
  return zeros;
}

}  // namespace pyr

extern "C" {

int pyr_deposit(pyr_mesh *mesh, double radius, double alpha, double normalizer, int precision,
                int64_t n, int vdim, const double *x, const double *q, const double *v,
                const uint32_t *elem, double *rho, double *j) {
  return guarded([&] {
    deposit_impl(mesh, radius, alpha, normalizer, precision, n, vdim, x, q, v, elem, rho, j);
  });
}

int pyr_deposit_normalized(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                           int precision, int64_t n, int vdim, const double *x, const double *q,
                           const double *v, const uint32_t *elem,
                           const double *integral_weights, double *rho, double *j,
                           double *scale_out, int64_t *n_zero_out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n_zero_out, "null argument");
    *n_zero_out = 0;
    PYR_REQUIRE(radius > 0 && std::isfinite(radius), "shape radius must be positive");
    PYR_REQUIRE(alpha > 0, "shape exponent must be positive");
    PYR_REQUIRE(precision == PYR_FP64 || precision == PYR_FP32C, "unknown precision");
    PYR_REQUIRE(vdim == mesh->dim, "vdim must equal the mesh dimension");
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    PYR_REQUIRE(!j || v || n == 0, "j requires velocities");
    PYR_REQUIRE(mesh->dev.jacobian, "the mesh was created without element jacobians");
    if (!rho && !j) return;
    if (n > 0) PYR_REQUIRE(x && q && elem && integral_weights, "null pointer");
    if (!mesh->dep) mesh->dep = new DepositCache();
    DepositCache *c = mesh->dep;
    const int K = mesh->K, Np = mesh->Np, D = mesh->dim;
    hipStream_t s = stream();
    const double *qn = q;
    const uint32_t *perm_kept = nullptr;
    if (n > 0) {
      prepare_tables(mesh, c, radius);
      if (c->tr_radius != radius) build_transpose(mesh, c);
      const uint32_t *perm;
      {
        ProfScope prof(PROF_SORT);
        perm = sort_and_stage(mesh, c, n, x, q, nullptr, elem, normalizer);
      }
      // FP64 shape integral per particle (dep_normshape.hpp:124-165)
      c->nrm.reserve(sizeof(double) * ((size_t)2 * n + 1) + sizeof(unsigned long long));
      double *integral = c->nrm.as<double>();
      double *qs = integral + n;
      unsigned long long *zeros = reinterpret_cast<unsigned long long *>(qs + n);
      PYR_HIP(hipMemsetAsync(integral, 0, sizeof(double) * (size_t)n, s));
      PYR_HIP(hipMemsetAsync(zeros, 0, sizeof(unsigned long long), s));
      IntegralArgs ia{};
      ia.K = K;
      ia.R = radius;
      ia.invR = 1.0 / radius;
      ia.alpha = alpha;
      ia.normalizer = normalizer;
      for (int a = 0; a < 3; ++a) ia.shift_len[a] = mesh->dev.per_max[a] - mesh->dev.per_min[a];
      ia.tr_start = c->tr_start.as<uint32_t>();
      ia.tr = c->tr.as<uint32_t>();
      ia.tr_shift = c->tr_shift.as<uint8_t>();
      ia.seg = c->seg.as<uint32_t>();
      ia.perm = perm;
      ia.pos = c->pos.as<double4>();
      ia.geom = c->geom.as<double>();
      ia.node_tab = c->tab64.as<double>();
      ia.iw = integral_weights;
      ia.jac = mesh->dev.jacobian;
      ia.integral = integral;
      const unsigned grid = build_chunks(mesh, c, n, ia.cl);
      {
        ProfScope prof(PROF_INTEGRAL);
        launch_shape_integral(D, Np, ia, grid, shape_unit_square(alpha, radius));
      }
      // charges q / I (dep_normshape.hpp:168-198); zero-mass particles drop out
      hipLaunchKernelGGL(k_normalized_charges, dim3(blocks_for(n)), dim3(256), 0, s, n, K, elem,
                         q, integral, qs, scale_out, zeros);
      PYR_CHECK_LAUNCH();
      unsigned long long hz = 0;
      PYR_HIP(hipMemcpyAsync(&hz, zeros, sizeof(hz), hipMemcpyDeviceToHost, s));
      PYR_HIP(hipStreamSynchronize(s));
      *n_zero_out = (int64_t)hz;
      qn = qs;
      perm_kept = perm;
    }
    // the deposit restages with the scaled charges on the same sort
    deposit_impl(mesh, radius, alpha, normalizer, precision, n, vdim, x, qn, v, elem, rho, j,
                 n > 0 ? perm_kept : nullptr);
  });
}

int pyr_average_forces(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                       int64_t n, int vdim, const double *const fields[6], const double *x,
                       const uint32_t *elem, const double *q, const double *v,
                       const double *integral_weights, double *force, double *integral_out,
                       double *vis_e, double *vis_b, double *vis_e_std, double *vis_b_std,
                       double *vis_el, double *vis_mag, int64_t *n_zero_out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n_zero_out && fields, "null argument");
    *n_zero_out = 0;
    double *vis[6] = {vis_e, vis_b, vis_e_std, vis_b_std, vis_el, vis_mag};
    unsigned long long *zeros =
        average_forces_impl(mesh, radius, alpha, normalizer, n, vdim, fields, x, elem, q, v,
                            integral_weights, force, integral_out, vis_e ? vis : nullptr);
    if (!zeros) return;
    unsigned long long hz = 0;
    PYR_HIP(hipMemcpyAsync(&hz, zeros, sizeof(hz), hipMemcpyDeviceToHost, stream()));
    PYR_HIP(hipStreamSynchronize(stream()));
    *n_zero_out = (int64_t)hz;
  });
}

int pyr_deposit_backward(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                         int precision, int64_t n, int vdim, const double *x, const double *q,
                         const double *v, const uint32_t *elem, const double *rho_bar,
                         const double *j_bar, double *x_bar, double *v_bar) {
  return guarded([&] {
    PYR_REQUIRE(mesh, "null mesh");
    PYR_REQUIRE(radius > 0 && std::isfinite(radius), "shape radius must be positive");
    PYR_REQUIRE(alpha >= 1, "the adjoint needs a shape exponent >= 1");
    PYR_REQUIRE(precision == PYR_FP64 || precision == PYR_FP32C, "unknown precision");
    PYR_REQUIRE(vdim == mesh->dim, "vdim must equal the mesh dimension");
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    if (n == 0) return;
    PYR_REQUIRE(!j_bar || (v && v_bar), "j_bar requires velocities and v_bar");
    PYR_REQUIRE(x && q && elem && x_bar, "null pointer");
    if (!mesh->dep) mesh->dep = new DepositCache();
    DepositCache *c = mesh->dep;
    const int K = mesh->K, Np = mesh->Np, D = mesh->dim;
    hipStream_t s = stream();
    PYR_HIP(hipMemsetAsync(x_bar, 0, sizeof(double) * (size_t)n * D, s));
    if (v_bar) PYR_HIP(hipMemsetAsync(v_bar, 0, sizeof(double) * (size_t)n * D, s));
    if (!rho_bar && !j_bar) return;
    prepare_tables(mesh, c, radius);
    if (c->tr_radius != radius) build_transpose(mesh, c);
    c->pos.reserve(sizeof(double4) * (size_t)n);
    c->wgt.reserve(sizeof(double4) * (size_t)n);
    const uint32_t *perm;
    {
      ProfScope prof(PROF_SORT);
      perm = sort_and_stage(mesh, c, n, x, q, j_bar ? v : nullptr, elem, normalizer);
    }
    AdjArgs ga{};
    ga.K = K;
    ga.R = radius;
    ga.invR = 1.0 / radius;
    ga.alpha = alpha;
    for (int a = 0; a < 3; ++a) ga.shift_len[a] = mesh->dev.per_max[a] - mesh->dev.per_min[a];
    ga.tr_start = c->tr_start.as<uint32_t>();
    ga.tr = c->tr.as<uint32_t>();
    ga.tr_shift = c->tr_shift.as<uint8_t>();
    ga.seg = c->seg.as<uint32_t>();
    ga.perm = perm;
    ga.pos = c->pos.as<double4>();
    ga.wgt = c->wgt.as<double4>();
    ga.geom = c->geom.as<double>();
    ga.node_tab = precision == PYR_FP64 ? (const void *)c->tab64.ptr : (const void *)c->tab32.ptr;
    ga.rho_bar = rho_bar;
    ga.j_bar = j_bar;
    ga.x_bar = x_bar;
    ga.v_bar = j_bar ? v_bar : nullptr;
    const unsigned grid = build_chunks(mesh, c, n, ga.cl);
    ProfScope prof(PROF_DEPOSIT_ADJ);
    launch_deposit_adj(D, Np, ga, grid, precision, alpha == 2.0);
  });
}

int pyr_deposit_set_counting(pyr_mesh *mesh, int on) {
  return guarded([&] {
    PYR_REQUIRE(mesh, "null mesh");
    mesh->deposit_count = on;
  });
}

int pyr_deposit_stats(pyr_mesh *mesh, unsigned long long out[4]) {
  return guarded([&] {
    PYR_REQUIRE(mesh && out, "null argument");
    for (int k = 0; k < 4; ++k) out[k] = 0;
    if (!mesh->dep || !mesh->dep->stats.ptr) return;
    PYR_HIP(hipStreamSynchronize(stream()));
    PYR_HIP(hipMemcpy(out, mesh->dep->stats.ptr, 4 * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost));
  });
}

int pyr_deposit_list_stats(pyr_mesh *mesh, unsigned long long out[4]) {
  return guarded([&] {
    PYR_REQUIRE(mesh && out, "null argument");
    for (int k = 0; k < 4; ++k) out[k] = 0;
    DepositCache *c = mesh->dep;
    if (!c || !c->list_valid) return;
    const int K = mesh->K;
    hipStream_t s = stream();
    c->list_stats.reserve(3 * sizeof(unsigned long long));
    PYR_HIP(hipMemsetAsync(c->list_stats.ptr, 0, 3 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_list_stats, dim3(blocks_for(K)), dim3(256), 0, s, K, mesh->dim,
                       c->cap.as<uint32_t>() + K + 1, c->n_live.as<uint32_t>(),
                       c->ent.as<uint2>(), c->list_stats.as<unsigned long long>());
    PYR_CHECK_LAUNCH();
    PYR_HIP(hipMemcpyAsync(out, c->list_stats.ptr, 3 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
    PYR_HIP(hipStreamSynchronize(s));
    out[3] = c->list_items;
  });
}

int pyr_deposit_candidates(pyr_mesh *mesh, double radius, int64_t *n_pairs_out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n_pairs_out && radius > 0, "invalid argument");
    if (!mesh->dep) mesh->dep = new DepositCache();
    if (mesh->dep->radius != radius) build_candidates(mesh, mesh->dep, radius);
    *n_pairs_out = mesh->dep->n_cand;
  });
}

int pyr_sort_by_element(pyr_mesh *mesh, int64_t n, const uint32_t *elem, uint32_t *perm_out,
                        uint32_t *seg_out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid argument");
    if (!mesh->dep) mesh->dep = new DepositCache();
    DepositCache *c = mesh->dep;
    if (n == 0) {
      if (seg_out)
        PYR_HIP(hipMemsetAsync(seg_out, 0, sizeof(uint32_t) * ((size_t)mesh->K + 1), stream()));
      return;
    }
    PYR_REQUIRE(elem && perm_out, "null pointer");
    const uint32_t *perm = sort_by_element(mesh, c, n, elem);
    PYR_HIP(hipMemcpyAsync(perm_out, perm, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice,
                           stream()));
    if (seg_out)
      PYR_HIP(hipMemcpyAsync(seg_out, c->seg.ptr, sizeof(uint32_t) * ((size_t)mesh->K + 1),
                             hipMemcpyDeviceToDevice, stream()));
  });
}

}  // extern "C"
