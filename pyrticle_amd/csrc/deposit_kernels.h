// deposit_kernels.h -- device code of the shape-function deposition, shared by
// the translation units that instantiate it: gather_fp64.hip (FP64 gather),
// gather_fp32.hip (FP32-compensated gather on the matrix cores) and
// deposit_adj.hip (adjoint).  Host code and the sort live in deposit.hip.
#pragma once

#include <type_traits>

#include "pyr_internal.h"

namespace pyr {

// ---------------------------------------------------------------------------
// gather kernel
// ---------------------------------------------------------------------------
struct GatherArgs {
  int K, Np;
  double R, invR, R2;
  double alpha;
  double shift_len[3];              // periodic extents (shift code 1: +L, 2: -L)
  const uint32_t *cand_start;       // K + 1
  const uint32_t *cand;             // source element ids
  const uint8_t *cand_shift;        // 2 bits per axis
  const uint32_t *seg;              // K + 1 segment starts into pos/wgt
  const double4 *pos;               // staged positions
  const double4 *wgt;               // staged weights
  const double *geom;               // K x 8: centroid (3), (R + node radius)^2, vertex radius
  const double *face_normal;        // mesh arrays
  const double *face_rhs;
  const void *node_tab;             // K x Np x 4 (T): n - c, |n - c|^2 / R
  double *rho;                      // K*Np or null
  double *j;                        // K*Np x D or null
  unsigned long long *stats;        // [pairs evaluated, pairs inside R, candidates, passed]
};

template <int D>
__device__ __forceinline__ void shift_vec(const GatherArgs &g, uint32_t code, double *s) {
#pragma unroll
  for (int a = 0; a < D; ++a) {
    uint32_t c = (code >> (2 * a)) & 3u;
    s[a] = c == 1 ? g.shift_len[a] : (c == 2 ? -g.shift_len[a] : 0.0);
  }
}

// max(t, 0)^2 as one multiply with an abs source and the clamp bit
// (v_mul_f64 / v_mul_f32 ... |t| clamp): clamp(t |t|) to [0, 1] equals
// max(t, 0)^2 bit for bit whenever t <= 1, and gives 0 for t = -inf.
template <class T>
__device__ __forceinline__ T sq_pos_unit(T t) {
  return fmin(fmax(t * fabs(t), T(0)), T(1));
}
// The FP64 shape kernels take that form for alpha = 2 when R <= 1/2: there
// t = R - |n - x|^2 / R <= R up to rounding, far below 1.  Otherwise they
// square (alpha = 2) or call pow, with the same bits as before.
inline bool shape_unit_square(double alpha, double radius) {
  return alpha == 2.0 && radius <= 0.5;
}

// Block -> element map.  The hardware deals consecutive blocks round-robin to
// the 8 XCDs; chunks of XCD_CHUNK consecutive elements (spatial neighbours that
// share candidate particles) go to one XCD and its L2, and the chunks rotate
// over the XCDs so that a dense region is spread over all of them.
#ifndef PYR_XCD_CHUNK
#define PYR_XCD_CHUNK 32
#endif
constexpr int XCD_CHUNK = PYR_XCD_CHUNK;
// (the FP64 matrix-core list gather takes runs of 128 blocks: its HBM traffic
// halves against runs of 32, 5.4 against 11.6 GB a launch on the 4M-particle
// cylinder, and it runs 2% faster; the FP32C one is no faster with them)
constexpr int XCD_CHUNK_MF64 = 128;
template <int C = XCD_CHUNK>
__device__ __forceinline__ int xcd_element(int bid) {
  const int xcd = bid & 7, slot = bid >> 3;
  const int chunk = (slot / C) * 8 + xcd;
  return chunk * C + slot % C;
}
template <int C = XCD_CHUNK>
static inline unsigned gather_grid(int K) {
  const int64_t per = 8 * C;
  return (unsigned)(((int64_t)K + per - 1) / per * per);
}

// 64-particle work chunks of the per-particle gathers: (source element, first
// sorted index); n_chunks (device) holds the count.
struct ChunkList {
  const uint2 *chunks;
  const uint32_t *n_chunks;
};

// The chunk a block of a per-particle gather handles (XCD-grouped: nearby
// chunks, and so nearby elements, share an L2), or false past the end.
__device__ __forceinline__ bool chunk_of(const ChunkList &cl, uint32_t &s, uint32_t &base) {
  const uint32_t k = (uint32_t)xcd_element(blockIdx.x);
  if (k >= *cl.n_chunks) return false;
  const uint2 ch = cl.chunks[k];
  s = ch.x;
  base = ch.y;
  return true;
}

// FP64 gather per target element.  A block holds G waves, each owning NPW of the
// element's nodes.  The block streams the candidate particles, drops those
// whose ball cannot reach the element (bounding sphere of the nodes inflated
// by R, face planes moved out by R), queues the rest in an LDS ring buffer and
// evaluates full batches of 64 queued particles one particle per lane against
// each wave's nodes (node data in scalar registers), accumulating rho and j
// per node in registers.
// A2: alpha = 2 with R <= 1/2 (shape_unit_square), squares by sq_pos_unit
template <int D, int NP, int NPW, bool A2, bool COUNT>
__global__ __launch_bounds__(64 * ((NP + NPW - 1) / NPW)) void k_gather(GatherArgs g) {
  using T = double;
  constexpr int G = (NP + NPW - 1) / NPW;   // waves per block = node groups
  constexpr int BS = 64 * G;
  constexpr int QCAP = 64 + BS;             // ring buffer of queued particles
  constexpr int NV = D + 1;                 // rho + j components
  constexpr int REC = 8;                    // record: m (3), c, a, b (3)

  const int e = xcd_element(blockIdx.x);
  if (e >= g.K) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform per wave
  const int node0 = wave * NPW;
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;

  __shared__ T q_rec[QCAP][REC];
  __shared__ int wave_pass[G];

  // target element geometry (uniform across the block)
  const double cx = g.geom[(size_t)e * 8 + 0], cy = g.geom[(size_t)e * 8 + 1],
               cz = g.geom[(size_t)e * 8 + 2], bound2 = g.geom[(size_t)e * 8 + 3];
  double fnl[D + 1][D], frl[D + 1];
#pragma unroll
  for (int f = 0; f <= D; ++f) {
#pragma unroll
    for (int k = 0; k < D; ++k) fnl[f][k] = g.face_normal[((size_t)e * (D + 1) + f) * D + k];
    frl[f] = g.face_rhs[(size_t)e * (D + 1) + f] + g.R * (1.0 + 1e-12);
  }
  // node data through the constant address space: uniform -> scalar loads
  using CT = const T __attribute__((address_space(4))) *;
  const CT ntab = (CT)(static_cast<const T *>(g.node_tab)) + (size_t)e * NP * 4;

  T acc[NPW][NV];
#pragma unroll
  for (int i = 0; i < NPW; ++i)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[i][v] = T(0);
  unsigned long long n_eval = 0, n_in = 0, n_cand = 0, n_pass = 0;
  int q_head = 0, q_count = 0;   // identical in every thread of the block
  bool any = false;

  auto process_batch = [&](int count) {
    // one queued particle per lane, all of this wave's nodes
    if (lane < count) {
      int slot = q_head + lane;
      if (slot >= QCAP) slot -= QCAP;
      T rec[REC];
#pragma unroll
      for (int k = 0; k < REC; ++k) rec[k] = q_rec[slot][k];
#pragma unroll
      for (int i = 0; i < NPW; ++i) {
        // compile-time true when the waves tile the nodes exactly, so that
        // the node loads are not split by branches
        if (NP % NPW == 0 || node0 + i < NP) {
          const CT nd = ntab + (size_t)(node0 + i) * 4;
          T t = rec[3] - nd[3];
#pragma unroll
          for (int k = D - 1; k >= 0; --k) t = fma(nd[k], rec[k], t);
          T s;
          if constexpr (A2) {
            s = sq_pos_unit(t);
          } else {
            const T w = t > T(0) ? t : T(0);
            s = g.alpha == 2.0 ? w * w : (T)pow((double)w, g.alpha);
          }
          if constexpr (COUNT) n_in += t > T(0) ? 1 : 0;
          acc[i][0] = fma(rec[4], s, acc[i][0]);
#pragma unroll
          for (int k = 0; k < D; ++k) acc[i][1 + k] = fma(rec[5 + k], s, acc[i][1 + k]);
        }
      }
      if constexpr (COUNT) n_eval += (unsigned long long)min(NPW, NP - node0);
    }
  };

  const uint32_t c_begin = g.cand_start[e], c_end = g.cand_start[e + 1];
  for (uint32_t cw = c_begin; cw < c_end; cw += 64) {
    // window of 64 candidate source elements: segment starts, lengths, shifts
    uint32_t len = 0, start = 0, sc = 0;
    if (cw + lane < c_end) {
      const uint32_t src = g.cand[cw + lane];
      start = g.seg[src];
      len = g.seg[src + 1] - start;
      sc = g.cand_shift[cw + lane];
    }
    uint64_t nonempty = __ballot(len > 0);
    while (nonempty) {
      const int k = __ffsll((unsigned long long)nonempty) - 1;
      nonempty &= nonempty - 1;
      const uint32_t s0 = __builtin_amdgcn_readlane(start, k);
      const uint32_t sl = __builtin_amdgcn_readlane(len, k);
      const uint32_t code = __builtin_amdgcn_readlane(sc, k);
      double sh[3] = {0, 0, 0};
      if (code) shift_vec<D>(g, code, sh);
      for (uint32_t base = 0; base < sl; base += BS) {
        const uint32_t idx = base + tid;
        bool pass = false;
        T rec[REC];
        if (idx < sl) {
          const double4 pp = g.pos[s0 + idx];
          const double x[3] = {pp.x + sh[0], pp.y + sh[1], D == 3 ? pp.z + sh[2] : 0.0};
          const double d[3] = {x[0] - cx, x[1] - cy, x[2] - cz};
          const double dd = fma(d[0], d[0], fma(d[1], d[1], D == 3 ? d[2] * d[2] : 0.0));
          // conservative: the ball of radius R misses the element if x lies
          // beyond one of its face planes by more than R
          double viol = -INFINITY;
#pragma unroll
          for (int f = 0; f <= D; ++f) {
            double ip = fnl[f][0] * x[0];
#pragma unroll
            for (int kk = 1; kk < D; ++kk) ip = fma(fnl[f][kk], x[kk], ip);
            viol = fmax(viol, ip - frl[f]);
          }
          pass = (dd <= bound2) & (viol <= 0.0);
          if constexpr (COUNT) n_cand += 1;
          if (pass) {
            if constexpr (COUNT) n_pass += 1;
            const double4 ww = g.wgt[s0 + idx];
            const double two_inv_r = 2.0 * g.invR;
            rec[0] = (T)(d[0] * two_inv_r);
            rec[1] = (T)(d[1] * two_inv_r);
            rec[2] = (T)(d[2] * two_inv_r);
            rec[3] = (T)fma(-dd, g.invR, g.R);
            rec[4] = (T)ww.x;
            rec[5] = (T)ww.y;
            rec[6] = (T)ww.z;
            rec[7] = (T)ww.w;
          }
        }
        const uint64_t bal = __ballot(pass);
        if (G > 1) {
          if (lane == 0) wave_pass[wave] = __popcll(bal);
          __syncthreads();
        }
        int off = q_count, added = 0;
        if (G > 1) {
#pragma unroll
          for (int w = 0; w < G; ++w) {
            if (w < wave) off += wave_pass[w];
            added += wave_pass[w];
          }
        } else {
          added = __popcll(bal);
        }
        if (pass) {
          int slot = q_head + off + __popcll(bal & lt_mask);
          if (slot >= QCAP) slot -= QCAP;
#pragma unroll
          for (int kk = 0; kk < REC; ++kk) q_rec[slot][kk] = rec[kk];
        }
        q_count += added;
        __syncthreads();
        while (q_count >= 64) {
          any = true;
          process_batch(64);
          q_head += 64;
          if (q_head >= QCAP) q_head -= QCAP;
          q_count -= 64;
          __syncthreads();   // the freed slots may be refilled next
        }
      }
    }
  }
  if (q_count > 0) {
    any = true;
    process_batch(q_count);
  }

  // reduce across lanes in a fixed tree and write this wave's nodes
#pragma unroll
  for (int i = 0; i < NPW; ++i) {
    const int node = node0 + i;
    if (node >= NP) continue;
    double out[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double val = any ? acc[i][v] : 0.0;
      if (any)
        for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, 64);
      out[v] = val;
    }
    if (lane == 0) {
      const size_t gn = (size_t)e * NP + node;
      if (g.rho) g.rho[gn] = out[0];
      if (g.j)
#pragma unroll
        for (int k = 0; k < D; ++k) g.j[gn * D + k] = out[1 + k];
    }
  }
  if constexpr (COUNT) {
    unsigned long long vals[4] = {n_eval, n_in, n_cand, n_pass};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      unsigned long long v = vals[k];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0 && v) atomicAdd(g.stats + k, v);
    }
  }
}

// ---------------------------------------------------------------------------
// FP32-compensated gather over culled chunk lists (matrix cores)
// ---------------------------------------------------------------------------
// Staging (k_stage32c, per deposition): the element-sorted particles of every
// source element are cut into chunks of up to 64 consecutive particles, 64
// slots each: pos = (x - c_src, 0) in FP32 (offset from the centroid of the
// particle's own element, formed in FP64 and rounded once), the weights
// w = (q N, q N v) of every aligned quad of particles stored transposed, and
// a bounding sphere of the chunk.  Empty slots carry zero weights.  The nodes
// of every target element are grouped spatially in groups of four
// (k_node_table_g).
//
// k_cull_chunks (one wave per target): for every chunk of every candidate
// source element, the groups with a node within R of the chunk's sphere;
// the chunks with at least one such group are written, in candidate order, to
// the target's list (chunk id and group mask; offset of the source centroid,
// + periodic shift, from the target centroid, formed in FP64 and rounded
// once).  A group none of whose nodes is within R of a chunk's particles adds
// exactly zero for that chunk, so the culling leaves every sum unchanged.
//
// k_gather_cl (one wave per work item of at most 64 consecutive entries of
// one target's list; items run in the order of their first chunk, so the
// waves in flight on an XCD share chunks in its L2): two pairs of chunks of
// loads in flight, each pair evaluated against the groups either chunk
// reaches, per group of four nodes n = 4 ng + (l % 4), on
// v_mfma_f32_4x4x1_16b_f32 (16 independent 4x4 outer products, block b =
// lanes 4b..4b+3):
//   t[p][n]  = (R - |n - x_p|^2 / R) / P,   P = 2^ceil(log2 R) >= R (exact scaling)
//            = -|n - c_e|^2 / (R P) + d_p . 2 (n - c_e) / (R P) + c_p,
//              d_p = x_p - c_e, c_p = (R - |d_p|^2 / R) / P (c_p enters
//              last: it carries the large cancelling part), as C + sum_k
//              A[p][k] B[k][n] over A = (d, c_p), B = (2 (n - c_e) / (R P), 1);
//              A of block b, row i is particle 4b + i = the lane's own
//              particle, so one instruction covers 64 particles x 4 nodes;
//              t lands in VGPR i of lane 4b + j: particle 4b + i, node j;
//   S        = max(t, 0)^alpha, one clamp to [0, 1] (t <= R / P <= 1; P^alpha
//              is folded into the staged weights);
//   acc[v][n] += sum_i wgtT_i S_i  (four K = 1 products, i = 0..3): the A
//              operand of block b is w_{4b+i} spread over its four rows, which
//              is component i of the lane's own transposed weight load.
// The two chunks of a pair accumulate on separate sets so their product chains
// interleave.  Every FOLD_STEPS chunks the per-lane FP32 partials are added
// to per-lane FP64 sums in LDS; at the end the sixteen blocks are combined in
// FP64 in a fixed butterfly and the item's sums written as FP64 partials,
// which k_combine_items adds per target in item order: results are bitwise
// reproducible.
typedef float pyr_v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ pyr_v4f mfma4x4(float a, float b, pyr_v4f c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

// chunks per fold into the FP64 sums (each FP32 partial then sums 4 x
// FOLD_STEPS particles per 4x4 block).  A work item holds at most 64 list
// entries, so its FP32 partials are folded once, at its end (on the 4M-beam
// cylinder the FP32C rho error against FP64 is 7.0e-8 with this, 6.9e-8 with
// a fold every 32 chunks)
constexpr int FOLD_STEPS = 64;

struct CullArgs {
  int K;
  double R;
  double shift_len[3];
  const uint32_t *cand_start;  // K + 1
  const uint32_t *cand;        // candidate source elements (ascending)
  const uint8_t *cand_shift;   // periodic shift codes
  const uint32_t *ch_start;    // K + 1: chunk range of each source element
  const float4 *ch_meta;       // per chunk: sphere centre rel. source centroid (3), radius
  const double *geom;          // K x 8: centroid (3), ...
  const float4 *ntab;          // K x 4 NG slots: 2 (n - c_e) / (R P) (3), -|n - c_e|^2 / (R P)
  const int *node_slot;        // 4 NG: node of each slot, -1 for padding
  float inv_rp;                // 1 / (R P)
  const uint32_t *cap_off;     // K + 1: list region of each target
  uint2 *ent;                  // list entries: chunk id, group mask | shift code << 16
  float4 *ent_off;             //               source offset from the target (3), 0
  uint32_t *n_live;            // K: list length
  const uint8_t *active;       // K: 0 when no candidate source element holds a particle
};

struct ListArgs {
  int K;
  double R, invR;
  float inv_p, alpha;          // 1 / P (P = 2^ceil(log2 R)), shape exponent
  const uint2 *items;          // work items: target, first list entry
  const uint32_t *item_order;  // processing order of the items
  const uint32_t *n_items;     // device count of work items
  const uint32_t *cap_off;     // K + 1
  const uint32_t *n_live;      // K
  const uint2 *ent;
  const float4 *ent_off;
  const float4 *pos;           // 64 slots per chunk: x - c_src (3), 0
  const float4 *wgt;           // 64 slots per chunk: weights of each quad, transposed
  const float4 *ntab;          // K x 4 NG slots: 2 (n - c_e) / (R P) (3), -|n - c_e|^2 / (R P)
  double *partial;             // per item: 4 NG slots x 4 components
  uint32_t zero_chunk;         // a staged chunk whose 64 slots are all zero
};

// relative margin of the FP32 sphere tests (they may only keep extra chunks)
constexpr float CL_MARGIN = 1e-4f;
// list entry word y: the group mask in the low bits (at most 9 groups), the
// periodic shift code of the chunk's source element from bit 16
constexpr int ENT_SHIFT_BIT = 16;
constexpr uint32_t ENT_MASK = (1u << ENT_SHIFT_BIT) - 1;
// work items (waves) per workgroup of the list gather (2 measured 3-4%
// faster than 4 on the 4M-particle cylinder, 1 the same as 2)
constexpr int CL_W = 2;
// word 3 of a raw buffer resource on gfx950 (32-bit data format; the untyped
// buffer loads ignore the format otherwise)
constexpr int BUF_WORD3 = 0x00020000;

template <int D, int NG>
__global__ __launch_bounds__(256) void k_cull_chunks(CullArgs g) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= g.K) return;   // uniform per wave
  if (!g.active[e]) {
    if (lane == 0) g.n_live[e] = 0;
    return;
  }
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  auto uni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
  const float rf = (float)g.R;
  // the nodes of every group from the gather's table, (2 n, -|n|^2) / (R P)
  // with n = node - c_e: a node lies within L of m when
  // (2 m . n - |n|^2) / (R P) >= (|m|^2 - L^2) / (R P) (the expanded form
  // rounds to ~1e-8 of the squared distances here, far inside the margin's
  // slack 2 CL_MARGIN (R + r)^2); padding slots never pass
  float nx[4 * NG], ny[4 * NG], nz[4 * NG], nn[4 * NG];
#pragma unroll
  for (int sl = 0; sl < 4 * NG; ++sl) {
    const float4 nd = g.ntab[(size_t)e * 4 * NG + sl];
    const bool pad = g.node_slot[sl] < 0;
    nx[sl] = uni(nd.x);
    ny[sl] = uni(nd.y);
    nz[sl] = uni(nd.z);
    nn[sl] = uni(pad ? -INFINITY : nd.w);
  }
  const float rn = rf * (1.f + CL_MARGIN);
  const double cx = g.geom[(size_t)e * 8 + 0], cy = g.geom[(size_t)e * 8 + 1],
               cz = D == 3 ? g.geom[(size_t)e * 8 + 2] : 0.0;
  const uint32_t base = g.cap_off[e];
  uint32_t n_out = 0;
  const uint32_t c_begin = g.cand_start[e], c_end = g.cand_start[e + 1];
  for (uint32_t cw = c_begin; cw < c_end; cw += 64) {
    uint32_t c0 = 0, cnt = 0, code = 0;
    float ox = 0.f, oy = 0.f, oz = 0.f;
    if (cw + lane < c_end) {
      const uint32_t src = g.cand[cw + lane];
      c0 = g.ch_start[src];
      cnt = g.ch_start[src + 1] - c0;
      if (cnt) {
        code = g.cand_shift[cw + lane];
        double sh[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const uint32_t c2 = (code >> (2 * a)) & 3u;
          sh[a] = c2 == 1 ? g.shift_len[a] : (c2 == 2 ? -g.shift_len[a] : 0.0);
        }
        ox = (float)(g.geom[(size_t)src * 8 + 0] + sh[0] - cx);
        oy = (float)(g.geom[(size_t)src * 8 + 1] + sh[1] - cy);
        oz = D == 3 ? (float)(g.geom[(size_t)src * 8 + 2] + sh[2] - cz) : 0.f;
      }
    }
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, 64);
      if (lane >= o) incl += y;
    }
    const uint32_t excl = incl - cnt;
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    for (uint32_t fb = 0; fb < total; fb += 64) {
      const uint32_t f = fb + (uint32_t)lane;
      int k = 0;   // the candidate lane holding flat chunk f
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const uint32_t v = __shfl(incl, k + step - 1, 64);
        if (v <= f) k += step;
      }
      k = min(k, 63);
      const uint32_t ch = __shfl(c0, k, 64) + (f - __shfl(excl, k, 64));
      const float px = __shfl(ox, k, 64), py = __shfl(oy, k, 64), pz = __shfl(oz, k, 64);
      // (in uniform control flow: a shuffle from a lane inactive in a branch
      // reads zero)
      const uint32_t pcode = (uint32_t)__shfl((int)code, k, 64);
      uint32_t ngm = 0;
      if (f < total) {
        const float4 meta = g.ch_meta[ch];
        const float mw = meta.w * (1.f + CL_MARGIN);
        const float mx = meta.x + px, my = meta.y + py, mz = D == 3 ? meta.z + pz : 0.f;
        const float ln = rn + mw;
        const float rhs = fmaf(-ln, ln, fmaf(mz, mz, fmaf(my, my, mx * mx))) * g.inv_rp;
        // a group is kept when one of its nodes lies within R of the chunk's
        // sphere (else every node is farther than R from every particle of
        // the chunk); all twenty tests run, without branches
#pragma unroll
        for (int ng = 0; ng < NG; ++ng) {
          bool near = false;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int sl = 4 * ng + i;
            near |= fmaf(nz[sl], mz, fmaf(ny[sl], my, fmaf(nx[sl], mx, nn[sl]))) >= rhs;
          }
          if (near) ngm |= 1u << ng;
        }
      }
      // the batch's kept chunks written grouped by group mask (in lane order
      // within a mask), so that the gather's pairs of consecutive entries
      // mostly share their groups
      uint64_t pend = __ballot(ngm != 0);
      while (pend) {
        const int l0 = __builtin_ctzll(pend);
        const uint32_t m0 = (uint32_t)__builtin_amdgcn_readlane((int)ngm, l0);
        const uint64_t grp = __ballot(ngm == m0) & pend;
        if ((grp >> lane) & 1ull) {
          const uint32_t at = base + n_out + (uint32_t)__popcll(grp & lt_mask);
          g.ent[at] = make_uint2(ch, ngm | (pcode << ENT_SHIFT_BIT));
          g.ent_off[at] = make_float4(px, py, pz, 0.f);
        }
        n_out += (uint32_t)__popcll(grp);
        pend &= ~grp;
      }
    }
  }
  if (lane == 0) g.n_live[e] = n_out;
}

template <int D, int NP, int W, bool A2>
__global__ __launch_bounds__(64 * W) void k_gather_cl(ListArgs g) {
  constexpr int NG = (NP + 3) / 4;
  constexpr int FOLD = FOLD_STEPS / 2;   // chunk pairs per fold
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t k_item = (int64_t)xcd_element(blockIdx.x) * W + wave;
  if (k_item >= (int64_t)*g.n_items) return;   // wave-uniform; no block-level synchronization
  const uint32_t item = g.item_order[k_item];
  const uint2 it = g.items[item];
  const uint32_t e = it.x;

  // FP64 sums of every accumulator entry per lane pair (l, l + 32); both
  // lanes of a pair hold the same value and write the same slot
  __shared__ double s_sum[W][NG * 4][32];
  double *kl = &s_sum[wave][0][lane & 31];
#pragma unroll
  for (int r = 0; r < NG * 4; ++r) kl[r * 32] = 0.0;

  // c_p = (R - |d|^2 / R) / P as one fma (P is a power of two, so the scaled
  // constants are the unscaled ones, exactly)
  const float r_p = (float)(g.R * (double)g.inv_p), inv_rp = (float)(g.invR * (double)g.inv_p);
  // this item's list entries: [first, first + cnt), one per lane
  const uint32_t base = g.cap_off[e] + it.y;
  const uint32_t cnt = min(64u, g.n_live[e] - it.y);

  // one accumulator set: the products of a pair's two chunks alternate on it
  // (a dependent 4x4x1 product takes its accumulator straight from the one
  // before)
  pyr_v4f acc[NG];
#pragma unroll
  for (int ng = 0; ng < NG; ++ng) acc[ng] = pyr_v4f{0.f, 0.f, 0.f, 0.f};
  int n_steps = 0;

  // lanes l and l + 32 are added in FP32 (v_permlane32_swap gives every lane
  // both, in the same order), then added to the pair's FP64 slot.  One group
  // at a time: the scheduler would otherwise hoist all the LDS reads of the
  // fold and hold them in registers.
  auto fold = [&]() {
#pragma unroll
    for (int ng = 0; ng < NG; ++ng) {
      double old[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) old[v] = kl[(ng * 4 + v) * 32];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t u = __float_as_uint(acc[ng][v]);
        const auto sw = __builtin_amdgcn_permlane32_swap(u, u, false, false);
        kl[(ng * 4 + v) * 32] = old[v] + (double)(__uint_as_float(sw[0]) +
                                                  __uint_as_float(sw[1]));
      }
      acc[ng] = pyr_v4f{0.f, 0.f, 0.f, 0.f};
      __builtin_amdgcn_sched_barrier(0);
    }
  };

  // node constants of this lane's node in every group
  float bx[NG], by[NG], bz[NG], cc[NG];
#pragma unroll
  for (int ng = 0; ng < NG; ++ng) {
    const float4 nt = g.ntab[(size_t)e * (4 * NG) + 4 * ng + (lane & 3)];
    bx[ng] = nt.x;
    by[ng] = nt.y;
    bz[ng] = nt.z;
    cc[ng] = nt.w;
  }
  // The item's list entries (chunk, group mask, source offset) are read with
  // scalar loads, four at a time one loop iteration before their chunks are
  // fetched; the chunk slots are read with buffer loads whose 64-bit base is
  // the chunk's (scalar) address and whose per-lane offset is fixed, so a
  // fetch costs no vector instruction besides its two loads.
  const uint32_t ubase = __builtin_amdgcn_readfirstlane(base);
  const auto s_ent = as_constant(reinterpret_cast<const uint32_t *>(g.ent + ubase));
  const auto s_off = as_constant(reinterpret_cast<const float *>(g.ent_off + ubase));
  const int lane_off = lane * (int)sizeof(float4);
  struct Ent {
    uint32_t ch, y;
    float ox, oy, oz, ow;
  };
  auto load_ent = [&](Ent &en, uint32_t i) {
    const uint32_t l = min(i, cnt - 1);
    en.ch = s_ent[2 * l];
    en.y = s_ent[2 * l + 1];
    en.ox = s_off[4 * l];
    en.oy = s_off[4 * l + 1];
    en.oz = s_off[4 * l + 2];
    en.ow = s_off[4 * l + 3];
  };
  struct Next {
    float4 p, w;
    float ox, oy, oz;
    uint32_t ngm;
  };
  // the loads of list entry i; past the end of the item (i >= cnt) the
  // all-zero chunk with no groups.  The loads are issued on every call, so
  // the same number is in flight on every path and the wait before a pair's
  // evaluation does not cover the pair fetched after it.
  auto fetch = [&](Next &n, uint32_t i, const Ent &en) {
    const bool live = i < cnt;   // wave-uniform
    const uint32_t ch = live ? en.ch : g.zero_chunk;
    // (the shift code above ENT_SHIFT_BIT is never tested: only the NG <= 9
    // group bits are, and every listed entry has one)
    n.ngm = live ? en.y : 0u;
    n.ox = en.ox;
    n.oy = en.oy;
    n.oz = en.oz;
    // the unused fourth word stays allocated until here: a register the
    // scalar load still writes is not handed out meanwhile (that reuse
    // would wait for the load)
    asm volatile("" ::"s"(en.ow));
    const size_t off = (size_t)ch * (64 * sizeof(float4));
    const auto rp = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(reinterpret_cast<const char *>(g.pos) + off), (short)0, 64 * (int)sizeof(float4),
        BUF_WORD3);
    const auto rw = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(reinterpret_cast<const char *>(g.wgt) + off), (short)0, 64 * (int)sizeof(float4),
        BUF_WORD3);
    const auto pv = __builtin_amdgcn_raw_buffer_load_b96(rp, lane_off, 0, 0);
    const auto wv = __builtin_amdgcn_raw_buffer_load_b128(rw, lane_off, 0, 0);
    n.p = make_float4(__uint_as_float(pv[0]), __uint_as_float(pv[1]), __uint_as_float(pv[2]), 0.f);
    n.w = make_float4(__uint_as_float(wv[0]), __uint_as_float(wv[1]), __uint_as_float(wv[2]),
                      __uint_as_float(wv[3]));
  };
  // t = (R - |n - x|^2 / R) / P: the chain starts from 1 * cc = cc exactly,
  // so it rounds as C = cc followed by the four products
  auto pair_t = [&](float dx, float dy, float dz, float cp, int ng) {
    pyr_v4f t = mfma4x4(1.f, cc[ng], pyr_v4f{0.f, 0.f, 0.f, 0.f});
    t = mfma4x4(dx, bx[ng], t);
    t = mfma4x4(dy, by[ng], t);
    if (D == 3) t = mfma4x4(dz, bz[ng], t);
    return mfma4x4(cp, 1.f, t);
  };
  auto shape = [&](float t) {
    if constexpr (A2) {
      // max(t, 0)^2 as clamp(t |t|) to [0, 1] (t <= 1 up to rounding)
      return __builtin_amdgcn_fmed3f(t * fabsf(t), 0.f, 1.f);
    } else {
      const float tw = __builtin_amdgcn_fmed3f(t, 0.f, 1.f);
      return tw > 0.f ? __powf(tw, g.alpha) : 0.f;
    }
  };
  // Two chunks against the groups either of them reaches.  A group the cull
  // rejected for one of them has t < 0 at every pair of that chunk (all its
  // nodes are farther than R from all its particles, by more than the FP32
  // rounding of t), so S = 0 there and its sums are unchanged.
  auto evaluate2 = [&](const Next &a, const Next &b) {
    const uint32_t mask = a.ngm | b.ngm;
    if (!mask) return;
    const float ax = a.p.x + a.ox, ay = a.p.y + a.oy, az = D == 3 ? a.p.z + a.oz : 0.f;
    const float acp = fmaf(-fmaf(az, az, fmaf(ay, ay, ax * ax)), inv_rp, r_p);
    const float bxx = b.p.x + b.ox, byy = b.p.y + b.oy, bzz = D == 3 ? b.p.z + b.oz : 0.f;
    const float bcp = fmaf(-fmaf(bzz, bzz, fmaf(byy, byy, bxx * bxx)), inv_rp, r_p);
#pragma unroll
    for (int ng = 0; ng < NG; ++ng) {
      if (!((mask >> ng) & 1u)) continue;
      const pyr_v4f ta = pair_t(ax, ay, az, acp, ng);
      const pyr_v4f tb = pair_t(bxx, byy, bzz, bcp, ng);
      acc[ng] = mfma4x4(a.w.x, shape(ta[0]), acc[ng]);
      acc[ng] = mfma4x4(b.w.x, shape(tb[0]), acc[ng]);
      acc[ng] = mfma4x4(a.w.y, shape(ta[1]), acc[ng]);
      acc[ng] = mfma4x4(b.w.y, shape(tb[1]), acc[ng]);
      acc[ng] = mfma4x4(a.w.z, shape(ta[2]), acc[ng]);
      acc[ng] = mfma4x4(b.w.z, shape(tb[2]), acc[ng]);
      acc[ng] = mfma4x4(a.w.w, shape(ta[3]), acc[ng]);
      acc[ng] = mfma4x4(b.w.w, shape(tb[3]), acc[ng]);
    }
    if (++n_steps == FOLD) {
      n_steps = 0;
      fold();
    }
  };
  // four named buffers: while one pair is evaluated the loads of the next
  // pair are in flight, and no buffer is copied.  The scheduling barriers
  // keep the loads of a fetch after the evaluation that frees its buffer (the
  // scheduler would otherwise start them early into fresh registers: 116
  // instead of 92 VGPRs, 4 instead of 5 waves per SIMD).
  Next a, b, c, d;
  Ent e0, e1, e2, e3;
  load_ent(e0, 0);
  load_ent(e1, 1);
  load_ent(e2, 2);
  load_ent(e3, 3);
  fetch(a, 0, e0);
  fetch(b, 1, e1);
  fetch(c, 2, e2);
  fetch(d, 3, e3);
#pragma unroll 1
  for (uint32_t i = 0; i < cnt; i += 4) {
    // the entries of this iteration's fetches, loaded one pair evaluation
    // before the first of them
    load_ent(e0, i + 4);
    load_ent(e1, i + 5);
    load_ent(e2, i + 6);
    load_ent(e3, i + 7);
    __builtin_amdgcn_sched_barrier(0);
    evaluate2(a, b);
    __builtin_amdgcn_sched_barrier(0);
    fetch(a, i + 4, e0);
    fetch(b, i + 5, e1);
    __builtin_amdgcn_sched_barrier(0);
    evaluate2(c, d);
    __builtin_amdgcn_sched_barrier(0);
    fetch(c, i + 6, e2);
    fetch(d, i + 7, e3);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (n_steps > 0) fold();
  // combine the sixteen blocks (slots of lanes l, l^4, ..., l^28, each
  // holding two blocks) in FP64 in a fixed butterfly; lanes 0..3 then hold
  // the item's sums of node slot 4 ng + lane, stored one by one (holding all
  // of them would set the register peak)
  double *dst = g.partial + (size_t)item * (16 * NG);
#pragma unroll
  for (int ng = 0; ng < NG; ++ng)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double val = kl[(ng * 4 + v) * 32];
#pragma unroll
      for (int o = 4; o < 32; o <<= 1) val += __shfl_xor(val, o, 64);
      if (lane < 4) dst[(4 * ng + lane) * 4 + v] = val;
    }
}

// ---------------------------------------------------------------------------
// FP64 gather over the culled chunk lists on the FP64 matrix cores
// ---------------------------------------------------------------------------
// v_mfma_f64_4x4x4_4b_f64 holds four blocks b (lane bits 2-3), each
// D = A B + C of 4x4 matrices with A[i][k] in lane 16 k + 4 b + i, B[k][j] in
// lane 16 k + 4 b + j and C, D[i][j] in lane 16 i + 4 b + j.  Per list entry
// (one 64-particle chunk, four slices of 16) and node group (four nodes j):
//   t  = C + A B with A[i][k] = (p_i (3), -|p_i|^2 / R), p = x - c_src (the
//        slice's particle 4 b + i, staged), B[k][j] = (b_j - 2 o / R, 1),
//        b = 2 (n_j - c_e) / R, o = c_src + shift - c_e, and
//        C[i][j] = R - |(n_j - c_e) - o|^2 / R, so that t is
//        R - |n_j - (x + shift)|^2 / R expanded around the source centroid
//        (C from the node's own offset to that centroid: the same cancellation
//        as the distance itself);
//   S  = max(t, 0)^alpha, elementwise: D's layout is the A layout of
//   acc[j][v] += sum_i S[i][j] w_i[v]  (A = S, B[i][v] = component v of the
//        block's particle i, staged transposed), one product per slice.
// Each block sums its own particles; the four blocks are added at the end in a
// fixed butterfly and the item's FP64 sums written as partials, which
// k_combine_items adds per target in item order.  A group an entry's mask
// leaves out has t < 0 for all of the chunk's particles and is skipped.
struct ListArgsM64 {
  int K;
  double R, invR, alpha;
  double shift_len[3];
  const uint2 *items;          // work items: target, first list entry
  const uint32_t *item_order;  // processing order of the items
  const uint32_t *n_items;     // device count of work items
  const uint32_t *cap_off;     // K + 1
  const uint32_t *n_live;      // K
  const uint2 *ent;            // list entries: chunk id, group mask | shift code << 16
  const uint2 *chunks;         // per chunk: source element, sorted index of its first particle
  const double *geom;          // K x 8: centroid (3), ...
  const double4 *node_tab;     // K x 4 NG slots: n - c_e (3), |n - c_e|^2 / R (padding: 2 R)
  const double *pos_t;         // 256 per chunk: the A operands of t, slice by slice
  const double *wgt_t;         // 256 per chunk: the B operands of the sums, slice by slice
  double *partial;             // per item: 4 NG slots x 4 components
};

__device__ __forceinline__ double mfma64(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

template <int D, int NP, int W, bool A2>
__global__ __launch_bounds__(64 * W) void k_gather_mf64(ListArgsM64 g) {
  constexpr int NG = (NP + 3) / 4;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t k_item = (int64_t)xcd_element<XCD_CHUNK_MF64>(blockIdx.x) * W + wave;
  if (k_item >= (int64_t)*g.n_items) return;   // wave-uniform
  const uint32_t item = __builtin_amdgcn_readfirstlane(g.item_order[k_item]);
  const uint2 it = g.items[item];
  const uint32_t e = __builtin_amdgcn_readfirstlane(it.x);
  const uint32_t first = __builtin_amdgcn_readfirstlane(it.y);
  const uint32_t base = __builtin_amdgcn_readfirstlane(g.cap_off[e] + first);
  const uint32_t cnt = __builtin_amdgcn_readfirstlane(min(64u, g.n_live[e] - first));
  const int kc = lane >> 4;   // row k of the lane's A and B operands of t
  const double R = g.R, invR = g.invR, m2r = -2.0 * g.invR;

  // the lane's node (column j = lane % 4) of every group: n - c_e (padding
  // slots at infinity: C = -inf, so t = -inf and S = 0 exactly) and its row
  // of B before the entry's shift (b[k] = 2 (n - c_e)[k] / R, 1 for k = 3)
  double nx[NG], ny[NG], nz[NG], bk[NG];
#pragma unroll
  for (int ng = 0; ng < NG; ++ng) {
    const double4 nt = g.node_tab[(size_t)e * (4 * NG) + 4 * ng + (lane & 3)];
    const bool pad = nt.w > 0.0 && nt.x == 0.0 && nt.y == 0.0 && nt.z == 0.0;
    nx[ng] = pad ? INFINITY : nt.x;
    ny[ng] = nt.y;
    nz[ng] = D == 3 ? nt.z : 0.0;
    bk[ng] = kc == 0 ? 2.0 * invR * nt.x
                     : (kc == 1 ? 2.0 * invR * nt.y : (kc == 2 ? 2.0 * invR * nz[ng] : 1.0));
  }
  const auto s_geom = as_constant(g.geom);
  const double cx = s_geom[(size_t)e * 8 + 0], cy = s_geom[(size_t)e * 8 + 1],
               cz = D == 3 ? s_geom[(size_t)e * 8 + 2] : 0.0;
  const auto s_ent = as_constant(reinterpret_cast<const uint32_t *>(g.ent + base));
  const auto s_chk = as_constant(reinterpret_cast<const uint32_t *>(g.chunks));

  double acc[NG];
#pragma unroll
  for (int ng = 0; ng < NG; ++ng) acc[ng] = 0.0;

  struct Entry {
    uint32_t mask;
    double ox, oy, oz;
    double p[4], w[4];
  };
  // entry i (clamped into the item: loads past the end are never evaluated)
  auto fetch = [&](Entry &en, uint32_t i) {
    const uint32_t l = min(i, cnt - 1);
    const uint32_t ch = s_ent[2 * l], y = s_ent[2 * l + 1];
    const uint32_t src = s_chk[2 * (size_t)ch];
    en.mask = y & ENT_MASK;
    const uint32_t code = y >> ENT_SHIFT_BIT;
    double sh[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t c2 = (code >> (2 * a)) & 3u;
      sh[a] = a < D ? (c2 == 1 ? g.shift_len[a] : (c2 == 2 ? -g.shift_len[a] : 0.0)) : 0.0;
    }
    en.ox = (s_geom[(size_t)src * 8 + 0] + sh[0]) - cx;
    en.oy = (s_geom[(size_t)src * 8 + 1] + sh[1]) - cy;
    en.oz = D == 3 ? (s_geom[(size_t)src * 8 + 2] + sh[2]) - cz : 0.0;
    const double *pp = g.pos_t + (size_t)ch * 256 + lane;
    const double *ww = g.wgt_t + (size_t)ch * 256 + lane;
#pragma unroll
    for (int sl = 0; sl < 4; ++sl) {
      en.p[sl] = pp[64 * sl];
      en.w[sl] = ww[64 * sl];
    }
  };
  auto shape = [&](double t) {
    if constexpr (A2) {
      return sq_pos_unit(t);
    } else {
      const double w = t > 0.0 ? t : 0.0;
      return g.alpha == 2.0 ? w * w : pow(w, g.alpha);
    }
  };
  auto evaluate = [&](const Entry &en) {
    const double ok = kc == 0 ? en.ox : (kc == 1 ? en.oy : (kc == 2 ? en.oz : 0.0));
#pragma unroll
    for (int ng = 0; ng < NG; ++ng) {
      if (!((en.mask >> ng) & 1u)) continue;
      const double b = fma(m2r, ok, bk[ng]);
      const double dx = nx[ng] - en.ox, dy = ny[ng] - en.oy, dz = nz[ng] - en.oz;
      const double c = fma(-fma(dz, dz, fma(dy, dy, dx * dx)), invR, R);
      double sv[4];
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) sv[sl] = shape(mfma64(en.p[sl], b, c));
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) acc[ng] = mfma64(sv[sl], en.w[sl], acc[ng]);
    }
  };

  // two named buffers: the loads of the next entry in flight while one is
  // evaluated
  Entry ea, eb;
  fetch(ea, 0);
#pragma unroll 1
  for (uint32_t i = 0; i < cnt; i += 2) {
    fetch(eb, i + 1);
    evaluate(ea);
    fetch(ea, i + 2);
    if (i + 1 < cnt) evaluate(eb);
  }

  // blocks (lane bits 2-3) added in a fixed butterfly; lanes 16 i + v then hold
  // node slot 4 ng + i, component v
  double *dst = g.partial + (size_t)item * (16 * NG);
#pragma unroll
  for (int ng = 0; ng < NG; ++ng) {
    double v = acc[ng];
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    if ((lane & 12) == 0) dst[(4 * ng + (lane >> 4)) * 4 + (lane & 3)] = v;
  }
}

// ---------------------------------------------------------------------------
// node skip of the per-particle gathers
// ---------------------------------------------------------------------------
// A sphere around the wave's valid particles (FP64: centre of their bounding
// box, radius to the farthest rounded up); a wave holds at least one.
__device__ __forceinline__ void wave_sphere(bool valid, const double4 &pp, double c[3],
                                            double &r) {
  double mn[3] = {valid ? pp.x : HUGE_VAL, valid ? pp.y : HUGE_VAL, valid ? pp.z : HUGE_VAL};
  double mx[3] = {valid ? pp.x : -HUGE_VAL, valid ? pp.y : -HUGE_VAL,
                  valid ? pp.z : -HUGE_VAL};
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fmin(mn[k], __shfl_xor(mn[k], o, 64));
      mx[k] = fmax(mx[k], __shfl_xor(mx[k], o, 64));
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = 0.5 * (mn[k] + mx[k]);
  const double dx = pp.x - c[0], dy = pp.y - c[1], dz = pp.z - c[2];
  double r2 = valid ? fma(dz, dz, fma(dy, dy, dx * dx)) : 0.0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
  r = sqrt(r2) * (1.0 + 1e-12) + 1e-300;
}

// The nodes of a target (rows of its node table: n - c_t, ...) that lie within
// lim of the point cc (relative to c_t): bit i for node i, one lane per node.
// A node outside has every particle of the wave's sphere farther than R from
// it when lim >= R + r (by the relative margin the callers add: 1e-9 over an
// FP64 table, 1e-5 over an FP32 one), so its shape value is exactly 0 for all
// of them and the wave can skip it.
template <int D, int NP, class T = double>
__device__ __forceinline__ uint64_t near_nodes(const T *nt, const double cc[3], double lim) {
  static_assert(NP <= 64, "one lane per node");
  const int lane = threadIdx.x & 63;
  bool near = false;
  if (lane < NP) {
    const double dx = (double)nt[lane * 4 + 0] - cc[0], dy = (double)nt[lane * 4 + 1] - cc[1],
                 dz = D == 3 ? (double)nt[lane * 4 + 2] - cc[2] : 0.0;
    near = fma(dz, dz, fma(dy, dy, dx * dx)) <= lim * lim;
  }
  return __ballot(near);
}

// ---------------------------------------------------------------------------
// adjoint of the deposition: per-particle gather of grad S . (rho_bar, j_bar)
// ---------------------------------------------------------------------------
// For particle p with a = q N and the nodes i within R:
//   rho_i = sum_p a S_ip,  j_ia = sum_p a v_pa S_ip,  S = w^alpha, w = R - |n_i - x_p|^2 / R
//   x_bar_p = (2/R) sum_i alpha w^(alpha-1) c_ip (n_i - x_p),  c_ip = a rho_bar_i + sum_a a v_pa j_bar_ia
//   v_bar_pa = a sum_i j_bar_ia S_ip
// Outputs are per particle, so a wave takes 64 particles of one source element
// and walks the target elements that element can reach (the transpose of the
// gather's candidate lists); target node data and adjoint densities are wave-
// uniform and come through scalar loads.  No atomics.
struct AdjArgs {
  ChunkList cl;
  int K;
  double R, invR, alpha;
  double shift_len[3];
  const uint32_t *tr_start;   // K + 1, per source element
  const uint32_t *tr;         // target elements
  const uint8_t *tr_shift;    // shift code applied to the source particle
  const uint32_t *seg;        // K + 1 segment starts into pos/wgt
  const uint32_t *perm;       // sorted position -> particle index
  const double4 *pos, *wgt;
  const double *geom;
  const void *node_tab;
  const double *rho_bar;      // K*Np or null
  const double *j_bar;        // K*Np x D or null
  double *x_bar, *v_bar;      // n x D (v_bar only with j_bar)
};

template <int D, int NP, bool FP32, bool A2>
__global__ __launch_bounds__(64) void k_deposit_adj(AdjArgs g) {
  using T = typename std::conditional<FP32, float, double>::type;
  using CT = const T __attribute__((address_space(4))) *;
  using CD = const double __attribute__((address_space(4))) *;
  uint32_t s, base;
  if (!chunk_of(g.cl, s, base)) return;
  const uint32_t p1 = g.seg[s + 1];
  const int lane = threadIdx.x;
  const uint32_t t_begin = g.tr_start[s], t_end = g.tr_start[s + 1];
  const double two_inv_r = 2.0 * g.invR;
  const CT ntab_all = (CT)(static_cast<const T *>(g.node_tab));
  const bool has_rho = g.rho_bar != nullptr, has_j = g.j_bar != nullptr;
  {
    const uint32_t jj = base + lane;
    const bool valid = jj < p1;
    double4 pp = make_double4(0, 0, 0, 0), ww = make_double4(0, 0, 0, 0);
    if (valid) {
      pp = g.pos[jj];
      ww = g.wgt[jj];
    }
    const T wa = (T)ww.x;
    const T wv[3] = {(T)ww.y, (T)ww.z, (T)ww.w};
    double wc[3], wr;
    wave_sphere(valid, pp, wc, wr);
    const double lim = g.R * (1.0 + (FP32 ? 1e-5 : 1e-9)) + wr;
    double gx[D], gv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      gx[k] = 0;
      gv[k] = 0;
    }
    for (uint32_t ti = t_begin; ti < t_end; ++ti) {
      const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.tr[ti]);
      const uint32_t code = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.tr_shift[ti]);
      double sh[3] = {0, 0, 0};
      if (code) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const uint32_t cc = (code >> (2 * a)) & 3u;
          sh[a] = cc == 1 ? g.shift_len[a] : (cc == 2 ? -g.shift_len[a] : 0.0);
        }
      }
      const CD gt = (CD)g.geom + (size_t)t * 8;
      const double d[3] = {pp.x + sh[0] - gt[0], pp.y + sh[1] - gt[1],
                           D == 3 ? pp.z + sh[2] - gt[2] : 0.0};
      const double dd = fma(d[0], d[0], fma(d[1], d[1], D == 3 ? d[2] * d[2] : 0.0));
      const bool pass = valid && dd <= gt[3];
      if (__ballot(pass) == 0) continue;
      const T m[3] = {(T)(d[0] * two_inv_r), (T)(d[1] * two_inv_r), (T)(d[2] * two_inv_r)};
      const T c0 = (T)fma(-dd, g.invR, g.R);
      const CT nt = ntab_all + (size_t)t * NP * 4;
      const double cc[3] = {wc[0] + sh[0] - gt[0], wc[1] + sh[1] - gt[1], wc[2] + sh[2] - gt[2]};
      const uint64_t nmask = near_nodes<D, NP, T>(static_cast<const T *>(g.node_tab) +
                                                      (size_t)t * NP * 4, cc, lim);
      const CD rb = (CD)g.rho_bar + (size_t)t * NP;
      const CD jb = (CD)g.j_bar + (size_t)t * NP * D;
      T sD = T(0), sDn[D], sS[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        sDn[k] = T(0);
        sS[k] = T(0);
      }
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        if (!((nmask >> i) & 1ull)) continue;   // uniform: w = 0 for the whole wave
        T tt = c0 - nt[i * 4 + 3];
#pragma unroll
        for (int k = D - 1; k >= 0; --k) tt = fma(nt[i * 4 + k], m[k], tt);
        const T w = (pass && tt > T(0)) ? tt : T(0);
        T cp = has_rho ? wa * (T)rb[i] : T(0);
        T jbi[D];
        if (has_j) {
#pragma unroll
          for (int a = 0; a < D; ++a) {
            jbi[a] = (T)jb[i * D + a];
            cp = fma(wv[a], jbi[a], cp);
          }
        }
        T dS, S;
        if constexpr (A2) {
          dS = w;   // dS/dw = 2w: the 2 is applied once at the end
          S = w * w;
        } else {
          dS = w > T(0) ? (T)(g.alpha * pow((double)w, g.alpha - 1.0)) : T(0);
          S = w > T(0) ? (T)pow((double)w, g.alpha) : T(0);
        }
        const T Dv = dS * cp;
        sD += Dv;
#pragma unroll
        for (int k = 0; k < D; ++k) sDn[k] = fma(Dv, nt[i * 4 + k], sDn[k]);
        if (has_j) {
#pragma unroll
          for (int a = 0; a < D; ++a) sS[a] = fma(jbi[a], S, sS[a]);
        }
      }
      // fold the target's partial sums in FP64: sum_i D_i (n_i - d)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        gx[k] += (double)sDn[k] - d[k] * (double)sD;
        gv[k] += (double)sS[k];
      }
    }
    if (valid) {
      const uint32_t idx = g.perm[jj];
      const double fx = two_inv_r * (A2 ? 2.0 : 1.0);
#pragma unroll
      for (int k = 0; k < D; ++k) g.x_bar[(size_t)idx * D + k] = fx * gx[k];
      if (g.v_bar) {
#pragma unroll
        for (int k = 0; k < D; ++k) g.v_bar[(size_t)idx * D + k] = ww.x * gv[k];
      }
    }
  }
}


// ---------------------------------------------------------------------------
// per-particle shape integral of the normalized deposition
// ---------------------------------------------------------------------------
// dep_normshape.hpp:124-165: I_p = sum_t |J_t| sum_i w_i S(n_ti - x_p) over the
// target elements t within reach of the particle (the transposed candidate
// lists of its source element), w = M 1 the integral weights of the reference
// element.  FP64 in both precisions (the charge normalization stays FP64).
// One wave per 64 particles of a source element; target data through scalar
// loads.  integral[perm[j]] receives I_p (including the shape normalizer).
struct IntegralArgs {
  ChunkList cl;
  int K;
  double R, invR, alpha, normalizer;
  double shift_len[3];
  const uint32_t *tr_start;   // K + 1, per source element
  const uint32_t *tr;         // target elements
  const uint8_t *tr_shift;    // shift code applied to the source particle
  const uint32_t *seg;        // K + 1 segment starts into pos
  const uint32_t *perm;       // sorted position -> particle index
  const double4 *pos;
  const double *geom;
  const double *node_tab;     // K x Np x 4 (FP64): n - c, |n - c|^2 / R
  const double *iw;           // Np integral weights
  const double *jac;          // K: |det| of the element maps
  double *integral;           // n
};

// A2: alpha = 2 with R <= 1/2 (shape_unit_square), squares by sq_pos_unit
template <int D, int NP, bool A2>
__global__ __launch_bounds__(64) void k_shape_integral(IntegralArgs g) {
  using CD = const double __attribute__((address_space(4))) *;
  uint32_t s, base;
  if (!chunk_of(g.cl, s, base)) return;
  const uint32_t p1 = g.seg[s + 1];
  const int lane = threadIdx.x;
  const uint32_t t_begin = g.tr_start[s], t_end = g.tr_start[s + 1];
  const double two_inv_r = 2.0 * g.invR;
  const CD ntab_all = (CD)g.node_tab, iw = (CD)g.iw, jac = (CD)g.jac;
  {
    const uint32_t jj = base + lane;
    const bool valid = jj < p1;
    const double4 pp = valid ? g.pos[jj] : make_double4(0, 0, 0, 0);
    double wc[3], wr;
    wave_sphere(valid, pp, wc, wr);
    const double lim = g.R * (1.0 + 1e-9) + wr;
    double integral = 0;
    for (uint32_t ti = t_begin; ti < t_end; ++ti) {
      const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.tr[ti]);
      const uint32_t code = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.tr_shift[ti]);
      double sh[3] = {0, 0, 0};
      if (code) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const uint32_t cc = (code >> (2 * a)) & 3u;
          sh[a] = cc == 1 ? g.shift_len[a] : (cc == 2 ? -g.shift_len[a] : 0.0);
        }
      }
      const CD gt = (CD)g.geom + (size_t)t * 8;
      const double d[3] = {pp.x + sh[0] - gt[0], pp.y + sh[1] - gt[1],
                           D == 3 ? pp.z + sh[2] - gt[2] : 0.0};
      const double dd = fma(d[0], d[0], fma(d[1], d[1], D == 3 ? d[2] * d[2] : 0.0));
      const bool pass = valid && dd <= gt[3];
      if (__ballot(pass) == 0) continue;
      const double m[3] = {d[0] * two_inv_r, d[1] * two_inv_r, d[2] * two_inv_r};
      // A2: a lane that does not pass gets c0 = -inf, so every t is -inf and
      // its square term 0
      const double c0 = A2 && !pass ? -HUGE_VAL : fma(-dd, g.invR, g.R);
      const CD nt = ntab_all + (size_t)t * NP * 4;
      const double cc[3] = {wc[0] + sh[0] - gt[0], wc[1] + sh[1] - gt[1], wc[2] + sh[2] - gt[2]};
      const uint64_t nmask = near_nodes<D, NP>(g.node_tab + (size_t)t * NP * 4, cc, lim);
      double el_integral = 0;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        if (!((nmask >> i) & 1ull)) continue;   // uniform: S = 0 for the whole wave
        double tt = c0 - nt[i * 4 + 3];
#pragma unroll
        for (int k = D - 1; k >= 0; --k) tt = fma(nt[i * 4 + k], m[k], tt);
        double sv;
        if constexpr (A2) {
          sv = sq_pos_unit(tt);
        } else {
          const double w = (pass && tt > 0.0) ? tt : 0.0;
          sv = w > 0.0 ? (g.alpha == 2.0 ? w * w : pow(w, g.alpha)) : 0.0;
        }
        el_integral = fma(sv, iw[i], el_integral);
      }
      integral = fma(el_integral, jac[t], integral);
    }
    if (valid) g.integral[g.perm[jj]] = g.normalizer * integral;
  }
}

// ---------------------------------------------------------------------------
// Shape-weighted field averages (push_average.hpp:38-443, the averaging
// pusher): per particle, over the same nodes the deposition reaches,
//   I   = sum_e |J_e| sum_i S_i w_i,
//   <F> = sum_e |J_e| sum_i F_i S_i w_i / I      (F: ex..ez, bx..bz),
// and the force q <E> + v x (q <B>), truncated to vdim.  The reference forms
// the same ratio with q S (or the normalized (q / I) S) as the deposit and
// divides by the reconstructed charge (force_averaging_target
// ::add_shape_on_element :135-168, el_/mag_force_averaging_target::end_particle
// :249-265, :309-330); the
// common factor cancels.  A particle whose I is 0 gets no force, zero
// averages and zero deviations.
// ---------------------------------------------------------------------------
struct AverageArgs {
  IntegralArgs b;                   // traversal; b.integral (may be null) receives N I
  const double *f[6];               // ex ey ez bx by bz, K x Np nodal; null = zero
  const double *q, *v;              // per particle (caller's order), v: n x vdim
  int vdim;
  double *force;                    // n x vdim
  double *vis_e, *vis_b;            // n x 3 averages, or null
  double *vis_e_std, *vis_b_std;    // n, or null
  double *vis_el, *vis_mag;         // n x vdim, or null
  unsigned long long *zeros;        // count of located particles with I = 0
};

// A2: alpha = 2 with R <= 1/2 (shape_unit_square), squares by sq_pos_unit
template <int D, int NP, bool A2, bool VIS>
__global__ __launch_bounds__(64) void k_shape_average(AverageArgs g) {
  using CD = const double __attribute__((address_space(4))) *;
  uint32_t s, base;
  if (!chunk_of(g.b.cl, s, base)) return;
  const uint32_t p1 = g.b.seg[s + 1];
  const int lane = threadIdx.x;
  const uint32_t t_begin = g.b.tr_start[s], t_end = g.b.tr_start[s + 1];
  const double two_inv_r = 2.0 * g.b.invR;
  const CD ntab_all = (CD)g.b.node_tab, iw = (CD)g.b.iw, jac = (CD)g.b.jac;
  {
    const uint32_t jj = base + lane;
    const bool valid = jj < p1;
    const double4 pp = valid ? g.b.pos[jj] : make_double4(0, 0, 0, 0);
    double wc[3], wr;
    wave_sphere(valid, pp, wc, wr);
    const double lim = g.b.R * (1.0 + 1e-9) + wr;
    double aI = 0, aF[6] = {0, 0, 0, 0, 0, 0}, aS[2] = {0, 0};
    for (uint32_t ti = t_begin; ti < t_end; ++ti) {
      const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.b.tr[ti]);
      const uint32_t code = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.b.tr_shift[ti]);
      double sh[3] = {0, 0, 0};
      if (code) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const uint32_t cc = (code >> (2 * a)) & 3u;
          sh[a] = cc == 1 ? g.b.shift_len[a] : (cc == 2 ? -g.b.shift_len[a] : 0.0);
        }
      }
      const CD gt = (CD)g.b.geom + (size_t)t * 8;
      const double d[3] = {pp.x + sh[0] - gt[0], pp.y + sh[1] - gt[1],
                           D == 3 ? pp.z + sh[2] - gt[2] : 0.0};
      const double dd = fma(d[0], d[0], fma(d[1], d[1], D == 3 ? d[2] * d[2] : 0.0));
      const bool pass = valid && dd <= gt[3];
      if (__ballot(pass) == 0) continue;
      const double m[3] = {d[0] * two_inv_r, d[1] * two_inv_r, d[2] * two_inv_r};
      // A2: a lane that does not pass gets c0 = -inf, so every t is -inf and
      // its square term 0
      const double c0 = A2 && !pass ? -HUGE_VAL : fma(-dd, g.b.invR, g.b.R);
      const CD nt = ntab_all + (size_t)t * NP * 4;
      const double cc[3] = {wc[0] + sh[0] - gt[0], wc[1] + sh[1] - gt[1], wc[2] + sh[2] - gt[2]};
      const uint64_t nmask = near_nodes<D, NP>(g.b.node_tab + (size_t)t * NP * 4, cc, lim);
      // c_i = S_i w_i for this element's nodes (0 for the skipped ones)
      double cw[NP];
      double eI = 0;
#pragma unroll
      for (int i = 0; i < NP; ++i) {
        cw[i] = 0.0;
        if (!((nmask >> i) & 1ull)) continue;   // uniform: S = 0 for the whole wave
        double tt = c0 - nt[i * 4 + 3];
#pragma unroll
        for (int k = D - 1; k >= 0; --k) tt = fma(nt[i * 4 + k], m[k], tt);
        double sv;
        if constexpr (A2) {
          sv = sq_pos_unit(tt);
        } else {
          const double w = (pass && tt > 0.0) ? tt : 0.0;
          sv = w > 0.0 ? (g.b.alpha == 2.0 ? w * w : pow(w, g.b.alpha)) : 0.0;
        }
        cw[i] = sv * iw[i];
        eI += cw[i];
      }
      const double J = jac[t];
      aI = fma(eI, J, aI);
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        if (!g.f[f]) continue;                      // uniform: a zero component
        const CD fv = (CD)g.f[f] + (size_t)t * NP;
        double eF = 0, eS = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
          eF = fma(fv[i], cw[i], eF);
          if (VIS) eS = fma(fv[i] * fv[i], cw[i], eS);
        }
        aF[f] = fma(eF, J, aF[f]);
        if (VIS) aS[f / 3] = fma(eS, J, aS[f / 3]);
      }
    }
    const bool zero = valid && aI == 0.0;
    if (zero) {
      const unsigned long long zb = __ballot(1);
      if (lane == __builtin_ffsll((long long)zb) - 1)
        atomicAdd(g.zeros, (unsigned long long)__popcll(zb));
    }
    if (!valid) return;
    const uint32_t pi = g.b.perm[jj];
    const double qv = g.q[pi];
    if (g.b.integral) g.b.integral[pi] = g.b.normalizer * aI;
    double avg[6] = {0, 0, 0, 0, 0, 0}, el[3] = {0, 0, 0}, mag[3] = {0, 0, 0};
    if (!zero) {
#pragma unroll
      for (int f = 0; f < 6; ++f) avg[f] = aF[f] / aI;
      double vv[3] = {0, 0, 0};
      for (int k = 0; k < g.vdim; ++k) vv[k] = g.v[(size_t)pi * g.vdim + k];
      double qb[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        el[k] = qv * avg[k];
        qb[k] = qv * avg[3 + k];
      }
      mag[0] = vv[1] * qb[2] - vv[2] * qb[1];
      mag[1] = vv[2] * qb[0] - vv[0] * qb[2];
      mag[2] = vv[0] * qb[1] - vv[1] * qb[0];
    }
    for (int k = 0; k < g.vdim; ++k) g.force[(size_t)pi * g.vdim + k] = el[k] + mag[k];
    if (VIS) {
      for (int k = 0; k < 3; ++k) {
        g.vis_e[(size_t)pi * 3 + k] = avg[k];
        g.vis_b[(size_t)pi * 3 + k] = avg[3 + k];
      }
      for (int k = 0; k < g.vdim; ++k) {
        g.vis_el[(size_t)pi * g.vdim + k] = el[k];
        g.vis_mag[(size_t)pi * g.vdim + k] = mag[k];
      }
      // push_average.hpp:190-212: sqrt(mean of |F|^2 - |mean F|^2), 0 if negative
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double sd = 0;
        if (!zero) {
          const double mean_sq = avg[3 * h] * avg[3 * h] + avg[3 * h + 1] * avg[3 * h + 1] +
                                 avg[3 * h + 2] * avg[3 * h + 2];
          const double var = aS[h] / aI - mean_sq;
          sd = var < 0 ? 0.0 : sqrt(var);
        }
        (h ? g.vis_b_std : g.vis_e_std)[pi] = sd;
      }
    }
  }
}

// Launchers, one per translation unit.  precision: PYR_FP64 or PYR_FP32C.
void launch_gather_fp64(int D, int Np, const GatherArgs &ga, int K, bool a2, bool count);
void launch_cull_chunks(int D, int Np, const CullArgs &ca);
void launch_gather_cl(int D, int Np, const ListArgs &la, uint32_t n_items, bool a2);
void launch_gather_mf64(int D, int Np, const ListArgsM64 &la, uint32_t n_items, bool a2);
void launch_deposit_adj(int D, int Np, const AdjArgs &ga, unsigned grid, int precision,
                        bool a2);
void launch_shape_integral(int D, int Np, const IntegralArgs &ga, unsigned grid, bool a2);
void launch_shape_average(int D, int Np, const AverageArgs &ga, unsigned grid, bool a2,
                          bool vis);

}  // namespace pyr
