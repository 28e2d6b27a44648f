// push.hip -- DG interpolation of E and B, the Lorentz force and the fused
// five-stage LSRK4 particle push with relocalization after every stage.
//
// Interpolation uses per-element monomial coefficients c = V^{-1} f, so that
// f(x) = sum_k c_k phi_k(r(x)) with phi_k = prod_i r_i^{e_ik}.  That is the
// interpolant the reference evaluates as sum_j l_j(x) f_j with Lagrange weights
// l = V^{-T} phi from an LU solve per particle (push_monomial.hpp:155-210,
// :99-111); here the solve is folded into the coefficients once per field set
// and each particle-stage costs Np products plus 6*Np FMAs.
#include <cmath>

#include "pyr_internal.h"

namespace pyr {

// ---------------------------------------------------------------------------
// monomial basis in node_tuples() order (graded, see discretization.py)
// ---------------------------------------------------------------------------
template <int D, int N>
struct Basis {
  static constexpr int NP = D == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
  int e[NP][3];
  constexpr Basis() : e{} {
    int k = 0;
    for (int total = 0; total <= N; ++total) {
      if (D == 2) {
        for (int j = 0; j <= total; ++j) {
          e[k][0] = total - j;
          e[k][1] = j;
          e[k][2] = 0;
          ++k;
        }
      } else {
        for (int j = 0; j <= total; ++j)
          for (int l = 0; l <= total - j; ++l) {
            e[k][0] = total - j - l;
            e[k][1] = j;
            e[k][2] = l;
            ++k;
          }
      }
    }
  }
};

template <int D, int N, class T>
__device__ __forceinline__ void monomials(const T *r, T *phi) {
  constexpr Basis<D, N> B{};
  T pw[D][N + 1];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    pw[i][0] = T(1);
#pragma unroll
    for (int e = 1; e <= N; ++e) pw[i][e] = pw[i][e - 1] * r[i];
  }
#pragma unroll
  for (int k = 0; k < B.NP; ++k) {
    T v = pw[0][B.e[k][0]];
#pragma unroll
    for (int i = 1; i < D; ++i) v = v * pw[i][B.e[k][i]];
    phi[k] = v;
  }
}

// Row stride of the modal coefficient tables [element][field][monomial]: the
// FP32 table pads every row to a multiple of four monomials (16-byte aligned
// rows for the matrix-core interpolation), the FP64 table is dense.
template <class T>
__host__ __device__ constexpr int coef_stride(int np) {
  return sizeof(T) == 4 ? (np + 3) & ~3 : np;
}

// six interpolated field values at unit coordinates r in element en
template <int D, int N, class T, class P>
__device__ __forceinline__ void interp6(P coef, uint32_t en, const T *r, T *f6) {
  constexpr int NP = Basis<D, N>::NP, NS = coef_stride<T>(NP);
  T phi[NP];
  monomials<D, N, T>(r, phi);
  const P c = coef + (size_t)en * 6 * NS;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < NP; ++k) acc = fma(c[f * NS + k], phi[k], acc);
    f6[f] = acc;
  }
}

// The same six values on the matrix cores, for a full wave (all 64 lanes
// active): f6^T = C^T Phi^T as Np rank-1 steps of v_mfma_f32_4x4x1_16b_f32.
// Block b (lanes 4b..4b+3) covers four fields (rows i, the A operand: lane
// 4b+i holds c[field i][k] of the block's element) for the block's four
// particles (columns j, the B operand: lane 4b+j holds its own phi_k), so VGPR
// i of every lane is field i at the lane's own particle; a second accumulator
// takes fields 4 and 5 (rows 2 and 3 zero).  A block whose particles sit in
// several elements takes one pass per element (the smallest pending element
// id of the block first), every lane keeping the pass of its own element; a
// wave runs as many passes as its most mixed block.  Each rank-1 step rounds
// like an FMA, in the same monomial order as the VALU contraction, so both give
// the same bits.
typedef float pyr_push_v4f __attribute__((ext_vector_type(4)));

// The element of every active lane, if the wave sits in one element (the
// common case for element-sorted particles), else INVALID.  A uniform element
// lets per-element data come through scalar loads instead of per-lane ones.
__device__ __forceinline__ uint32_t wave_uniform_element(uint32_t e) {
  const uint32_t e0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)e);
  return __ballot(e != e0) == 0 ? e0 : INVALID;
}

template <int D, int N, class T>
__device__ __forceinline__ void interp6_any(const T *__restrict__ coef, uint32_t e,
                                            uint32_t e_uni, const T *r, T *f6) {
  if (e_uni != INVALID)
    interp6<D, N, T>(as_constant(coef), (uint32_t)__builtin_amdgcn_readfirstlane((int)e_uni),
                     r, f6);
  else
    interp6<D, N, T>(coef, e, r, f6);
}

// push_monomial.hpp:265-305 with tools.hpp:253-264 cross: F = q e + v x (q b),
// the velocity padded with zeros to three components (entry_or_zero).
template <int VD, class T>
__device__ __forceinline__ void lorentz(T q, const T *f6, const T *v, T *force, T *el,
                                        T *mag) {
  T qb[3], vv[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    el[i] = q * f6[i];
    qb[i] = q * f6[3 + i];
    vv[i] = i < VD ? v[i] : T(0);
  }
  mag[0] = vv[1] * qb[2] - vv[2] * qb[1];
  mag[1] = vv[2] * qb[0] - vv[0] * qb[2];
  mag[2] = vv[0] * qb[1] - vv[1] * qb[0];
#pragma unroll
  for (int i = 0; i < VD; ++i) force[i] = el[i] + mag[i];
}

// ---------------------------------------------------------------------------
// nodal -> modal coefficients: out[e][f][k] = sum_j Minv[k][j] field_f[e*Np+j]
// ---------------------------------------------------------------------------
struct Fields6 {
  const double *f[6];
};

// (rows of coef_stride<T>(np) entries; the padding is zero).  With occ, only
// the rows of elements e with occ[e] == stamp are written.
template <class T>
__global__ void k_modal(int K, int np, Fields6 fl, const double *__restrict__ minv,
                        T *__restrict__ out, const uint32_t *__restrict__ occ, uint32_t stamp) {
  const int ns = coef_stride<T>(np);
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)K * ns) return;
  int e = (int)(t / ns), k = (int)(t % ns);
  if (occ && occ[e] != stamp) return;
  for (int f = 0; f < 6; ++f) {
    double acc = 0;
    if (fl.f[f] && k < np) {
      const double *src = fl.f[f] + (size_t)e * np;
      for (int j = 0; j < np; ++j) acc = fma(minv[k * np + j], src[j], acc);
    }
    out[((size_t)e * 6 + f) * ns + k] = (T)acc;
  }
}

// ---------------------------------------------------------------------------
// drop-in forces (push_monomial.hpp:219-325)
// ---------------------------------------------------------------------------
template <int D, int N>
__global__ __launch_bounds__(256) void k_forces(DevMesh m, const double *__restrict__ coef, int64_t n,
                         const double *__restrict__ x, const uint32_t *__restrict__ elem,
                         const double *__restrict__ q, const double *__restrict__ v,
                         double *__restrict__ force, double *vis_e, double *vis_b,
                         double *vis_el, double *vis_mag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[D], r[D], f6[6], vi[D], F[D], el[3], mag[3];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xi[k] = x[i * D + k];
    vi[k] = v[i * D + k];
  }
  const uint32_t en = elem[i];
  if (en >= (uint32_t)m.K) {   // not located: the reference has no defined result
#pragma unroll
    for (int k = 0; k < D; ++k) force[i * D + k] = __builtin_nan("");
    return;
  }
  inverse_map<D>(m, en, xi, r);
  interp6<D, N, double>(coef, en, r, f6);
  lorentz<D, double>(q[i], f6, vi, F, el, mag);
#pragma unroll
  for (int k = 0; k < D; ++k) force[i * D + k] = F[k];
  if (vis_e) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      vis_e[i * 3 + k] = f6[k];
      vis_b[i * 3 + k] = f6[3 + k];
      vis_el[i * 3 + k] = el[k];
      vis_mag[i * 3 + k] = mag[k];
    }
  }
}

// ---------------------------------------------------------------------------
// fused LSRK4 step (hedge LSRK4TimeStepper, Carpenter & Kennedy 1994)
// ---------------------------------------------------------------------------
__constant__ double c_rk4a[5] = {
    0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
    -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
__constant__ double c_rk4b[5] = {
    1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
    1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
    2277821191437.0 / 14882151754819.0};

// The update half of a stage for one particle: k = [v(p), F(x, p)] at unit
// coordinates r in element e, res = a_s res + dt k, y += b_s res.  Returns
// the wave-uniform element (INVALID if the wave's particles are in several).
template <int D, int N, class T>
__device__ __forceinline__ uint32_t stage_update(const T *__restrict__ coef, int s, double dt,
                                                 double c, double qi, double mi, double *xi,
                                                 double *pi, double *rx, double *rp,
                                                 const double *r, uint32_t e) {
  double v[D], F[D];
  velocity<D>(pi, mi, c, v);
  const uint32_t e_uni = wave_uniform_element(e);
  if constexpr (sizeof(T) == 8) {
    double f6[6], el[3], mag[3];
    interp6_any<D, N, double>(coef, e, e_uni, r, f6);
    lorentz<D, double>(qi, f6, v, F, el, mag);
  } else {
    float rf[D], f6[6], vf[D], Ff[D], el[3], mag[3];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      rf[k] = (float)r[k];
      vf[k] = (float)v[k];
    }
    interp6_any<D, N, float>(coef, e, e_uni, rf, f6);
    lorentz<D, float>((float)qi, f6, vf, Ff, el, mag);
#pragma unroll
    for (int k = 0; k < D; ++k) F[k] = (double)Ff[k];
  }
  const double a = c_rk4a[s], b = c_rk4b[s];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    rx[k] = a * rx[k] + dt * v[k];
    rp[k] = a * rp[k] + dt * F[k];
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xi[k] += b * rx[k];
    pi[k] += b * rp[k];
  }
  return e_uni;
}

// Unit coordinates of x in element e (the first test of
// find_new_containing_element): through scalar loads when e is wave-uniform,
// the same arithmetic either way.  True if x is still inside e.
template <int D>
__device__ __forceinline__ bool still_inside(const DevMesh &m, uint32_t e, uint32_t e_uni,
                                             const double *xi, double *r) {
  if (e_uni != INVALID) {
    const uint32_t eu = (uint32_t)__builtin_amdgcn_readfirstlane((int)e_uni);
    const_ptr<double> am = as_constant(m.inv_map) + (size_t)eu * (D * D + D);
#pragma unroll
    for (int ii = 0; ii < D; ++ii) {
      double t = 0;
#pragma unroll
      for (int jj = 0; jj < D; ++jj) t += am[ii * D + jj] * xi[jj];
      r[ii] = t + am[D * D + ii];
    }
    return in_unit_simplex<D>(r);
  }
  return in_element<D>(m, e, xi, r);
}

// One stage of the fused step for one particle: the update, then
// relocalization.  Returns false if the particle left the mesh.  Every forward
// variant (the fast kernel, the per-step kernel, the checkpointed one and the
// adjoint's recompute) runs stage_update and the same first containment test,
// so they agree bitwise.
template <int D, int N, class T>
__device__ __forceinline__ bool lsrk4_stage(const DevMesh &m, const T *__restrict__ coef, int s,
                                            double dt, double c, double qi, double mi,
                                            double *xi, double *pi, double *rx, double *rp,
                                            double *r, uint32_t &e, FindCounters &cnt,
                                            int &e_flag) {
  const uint32_t e_uni = stage_update<D, N, T>(coef, s, dt, c, qi, mi, xi, pi, rx, rp, r, e);
  // still inside the (wave-uniform) element: the first branch of
  // find_new_containing_element through scalar loads; otherwise the cascade,
  // with every lane of the wave helping the fan and grid searches
  bool in = false;
  if (e_uni != INVALID) {
    in = still_inside<D>(m, e, e_uni, xi, r);
    if (in) cnt.same++;
  }
  return relocate_wave<D>(m, !in, e_uni != INVALID, xi, pi, e, cnt, e_flag, r);
}

// Minimum waves per SIMD the push is compiled for (a register budget of
// 512 / waves VGPRs per lane).  The FP64 push at order >= 4 would otherwise
// take 264 VGPRs (1 wave per SIMD); capped at 168 (3 waves) it spills a few
// registers and runs 24% faster on the dense P4 workload (9.6 -> 7.3 ms).
// The other forms fit their occupancy without a cap.
template <int N, class T>
constexpr int push_min_waves() {
  return (sizeof(T) == 8 && N >= 4) ? 3 : 1;
}
// The cached fused step: the FP32C form at order <= 3 fits 4 waves per SIMD
// (128 VGPRs) without spills and runs 10% faster there (4M-particle cylinder,
// 0.70 -> 0.63 ms); the FP64 form would spill (0.73 -> 0.82 ms) and keeps
// push_min_waves.
template <int N, class T>
constexpr int step_min_waves() {
  return (sizeof(T) == 4 && N <= 3) ? 4 : push_min_waves<N, T>();
}

// ---------------------------------------------------------------------------
// fused LSRK4 step with a per-wave LDS element cache
// ---------------------------------------------------------------------------
// A wave's particles start a step in one or two elements (the deposition's
// sort order) and spread into their face neighbours as they cross faces; after
// the first stage three quarters of the waves hold several elements (2.1 on
// average on the 4M-particle cylinder).  Per-lane loads of the field
// coefficients, inverse maps and face data then cost the texture data path a
// full wave of data per load instruction, which saturated it (TD busy 87-91%
// of the push).  Here each wave copies, once per step and with coalesced
// loads, the data of its start elements and of their face neighbours into its
// own LDS slots, and the five stages read them from LDS: interpolation
// (6 x Np coefficients), the containment test (inverse map) and the first
// steps of relocation (face normals, the neighbour across the chosen face and
// its inverse map).  Element-uniform waves keep the scalar loads of the
// uncached kernel.  A lane in an element without a slot (two faces crossed in
// one step, the vertex fan, the global search, a periodic trip) sends its wave
// through the uncached stage until the lane is back in a cached element.  The
// arithmetic is that of lsrk4_stage on the same values: the same bits.
template <int D, int N, class T>
struct StepCache {
  static constexpr int NP = Basis<D, N>::NP, NS = coef_stride<T>(NP);
  static constexpr int INV = 6 * NS * (int)sizeof(T);    // byte offsets in a slot
  static constexpr int NRM = INV + (D * D + D) * 8;
  static constexpr int NBR = NRM + (D + 1) * D * 8;      // neighbour ids (u32)
  static constexpr int NBS = NBR + (D + 1) * 4;          // neighbour slots (int)
  static constexpr int BYTES = (NBS + (D + 1) * 4 + 15) / 16 * 16;
  static constexpr int Q = NBR / 16;                     // 16-byte chunks copied per slot
  // about 10 KB per wave: 4 waves per SIMD
  static constexpr int SLOTS = 10240 / BYTES < 16 ? 10240 / BYTES : 16;
  static_assert(INV % 16 == 0 && NRM % 16 == 0 && NBR % 16 == 0, "slot layout");
  static_assert(SLOTS >= D + 2, "a slot set must hold one element and its neighbours");
};

// The six values on the matrix cores from the cached coefficients, for a
// full wave: the rank-1 form above, block b's rows taken from the slot of the
// pass (the smallest pending slot of its four lanes), every lane keeping the
// pass of its own slot.
template <int D, int N>
__device__ __forceinline__ void interp6_mfma_slots(const char *base, int slot, const float *r,
                                                   float *f6) {
  using C = StepCache<D, N, float>;
  constexpr int NP = Basis<D, N>::NP, NS = coef_stride<float>(NP);
  float phi[NS];
  monomials<D, N, float>(r, phi);
#pragma unroll
  for (int k = NP; k < NS; ++k) phi[k] = 0.f;
  const int row = threadIdx.x & 3;
  const bool rb = row < 2;
  constexpr int NONE = 1 << 30;
  bool done = false;
  for (;;) {
    int cs = done ? NONE : slot;
    cs = min(cs, __shfl_xor(cs, 1, 64));
    cs = min(cs, __shfl_xor(cs, 2, 64));
    if (__ballot(cs != NONE) == 0) break;
    const int sl = cs == NONE ? 0 : cs;
    const float4 *ca = reinterpret_cast<const float4 *>(base + sl * C::BYTES) + row * (NS / 4);
    const float4 *cb =
        reinterpret_cast<const float4 *>(base + sl * C::BYTES) + (4 + (row & 1)) * (NS / 4);
    pyr_push_v4f a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k4 = 0; k4 < NS / 4; ++k4) {
      const float4 x = ca[k4];
      const float4 y = rb ? cb[k4] : make_float4(0.f, 0.f, 0.f, 0.f);
      a = __builtin_amdgcn_mfma_f32_4x4x1f32(x.x, phi[4 * k4 + 0], a, 0, 0, 0);
      b = __builtin_amdgcn_mfma_f32_4x4x1f32(y.x, phi[4 * k4 + 0], b, 0, 0, 0);
      a = __builtin_amdgcn_mfma_f32_4x4x1f32(x.y, phi[4 * k4 + 1], a, 0, 0, 0);
      b = __builtin_amdgcn_mfma_f32_4x4x1f32(y.y, phi[4 * k4 + 1], b, 0, 0, 0);
      a = __builtin_amdgcn_mfma_f32_4x4x1f32(x.z, phi[4 * k4 + 2], a, 0, 0, 0);
      b = __builtin_amdgcn_mfma_f32_4x4x1f32(y.z, phi[4 * k4 + 2], b, 0, 0, 0);
      a = __builtin_amdgcn_mfma_f32_4x4x1f32(x.w, phi[4 * k4 + 3], a, 0, 0, 0);
      b = __builtin_amdgcn_mfma_f32_4x4x1f32(y.w, phi[4 * k4 + 3], b, 0, 0, 0);
    }
    if (!done && cs == slot) {
      f6[0] = a[0];
      f6[1] = a[1];
      f6[2] = a[2];
      f6[3] = a[3];
      f6[4] = b[0];
      f6[5] = b[1];
      done = true;
    }
  }
}

template <int D, int N, class T>
__device__ __forceinline__ int cache_lookup(const uint32_t *cel, int nsl, uint32_t e) {
  int s = -1;
  for (int j = 0; j < nsl; ++j)
    if (cel[j] == e) {
      s = j;
      break;
    }
  return s;
}

// the k-th set bit of mask (k < popcount(mask))
__device__ __forceinline__ int nth_bit(uint32_t mask, int k) {
  for (int j = 0; j < k; ++j) mask &= mask - 1;
  return __builtin_ctz(mask);
}

// Chunk c (16 bytes) of the copy of the slots in mask: its source (the slot's
// element's coefficients, inverse map and face normals lie in a slot as in
// their tables) and its byte offset in the wave's slots.
template <int D, int N, class T>
__device__ __forceinline__ uint4 cache_chunk(const DevMesh &m, const T *__restrict__ coef,
                                             const uint32_t *cel, uint32_t mask, int c,
                                             int &dst) {
  using C = StepCache<D, N, T>;
  const int s = nth_bit(mask, c / C::Q), qq = c % C::Q;
  const uint32_t el = cel[s];
  const char *src;
  if (qq * 16 < C::INV)
    src = (const char *)coef + (size_t)el * C::INV + qq * 16;
  else if (qq * 16 < C::NRM)
    src = (const char *)m.inv_map + (size_t)el * (D * D + D) * 8 + (qq * 16 - C::INV);
  else
    src = (const char *)m.face_normal + (size_t)el * (D + 1) * D * 8 + (qq * 16 - C::NRM);
  dst = s * C::BYTES + qq * 16;
  return *(const uint4 *)src;
}

// Gives slots to the elements of the active lanes in `need` (lowest lane's
// element first), then to the face neighbours of those elements, in slots no
// lane holds (used), as many as fit; copies their data with the active lanes
// (four 16-byte chunks in flight per lane) and rewrites every slot's
// neighbour slots.  Lanes whose element finds no slot keep slot = -1.
template <int D, int N, class T>
__device__ __forceinline__ void cache_update(const DevMesh &m, const T *__restrict__ coef,
                                             char *base, uint32_t *cel, int &nsl, uint32_t used,
                                             uint64_t need, uint32_t e, int &slot) {
  using C = StepCache<D, N, T>;
  constexpr uint32_t ALL = C::SLOTS >= 32 ? ~0u : (1u << C::SLOTS) - 1;
  uint32_t fresh = 0;
  while (need && (used & ALL) != ALL) {
    const int j = __builtin_ctz(~used & ALL);
    const uint32_t eu = (uint32_t)__builtin_amdgcn_readlane((int)e, (int)__builtin_ctzll(need));
    cel[j] = eu;
    used |= 1u << j;
    fresh |= 1u << j;
    nsl = j + 1 > nsl ? j + 1 : nsl;
    const bool mine = e == eu;
    if (mine) slot = j;
    need &= ~__ballot(mine);
  }
  wave_lds_fence();
  uint32_t fr = fresh;
  while (fr && (used & ALL) != ALL) {
    const int j0 = __builtin_ctz(fr);
    fr &= fr - 1;
    const uint32_t ej = (uint32_t)__builtin_amdgcn_readfirstlane((int)cel[j0]);
#pragma unroll
    for (int f = 0; f <= D; ++f) {
      const uint32_t nb = as_constant(m.face_neighbor)[(size_t)ej * (D + 1) + f];
      if (nb == INVALID || (used & ALL) == ALL) continue;
      const int have = __builtin_amdgcn_readfirstlane(cache_lookup<D, N, T>(cel, nsl, nb));
      if (have >= 0) {
        used |= 1u << have;   // keep it
        continue;
      }
      const int j = __builtin_ctz(~used & ALL);
      cel[j] = nb;
      used |= 1u << j;
      fresh |= 1u << j;
      nsl = j + 1 > nsl ? j + 1 : nsl;
      wave_lds_fence();
    }
  }
  wave_lds_fence();
  const uint64_t act = __ballot(1);
  const int nact = __popcll(act);
  const int rank = __popcll(act & ((1ull << (threadIdx.x & 63)) - 1));
  const int total = __builtin_popcount(fresh) * C::Q;
  for (int c0 = rank; c0 < total; c0 += 4 * nact) {
    const int cb = c0 + nact, ccc = c0 + 2 * nact, cd = c0 + 3 * nact;
    int da = 0, db = 0, dc = 0, dd = 0;
    uint4 va = {}, vb = {}, vc = {}, vd = {};
    va = cache_chunk<D, N, T>(m, coef, cel, fresh, c0, da);
    if (cb < total) vb = cache_chunk<D, N, T>(m, coef, cel, fresh, cb, db);
    if (ccc < total) vc = cache_chunk<D, N, T>(m, coef, cel, fresh, ccc, dc);
    if (cd < total) vd = cache_chunk<D, N, T>(m, coef, cel, fresh, cd, dd);
    *(uint4 *)(base + da) = va;
    if (cb < total) *(uint4 *)(base + db) = vb;
    if (ccc < total) *(uint4 *)(base + dc) = vc;
    if (cd < total) *(uint4 *)(base + dd) = vd;
  }
  for (int w = rank; w < __builtin_popcount(fresh) * (D + 1); w += nact) {
    const int s = nth_bit(fresh, w / (D + 1)), f = w % (D + 1);
    ((uint32_t *)(base + s * C::BYTES + C::NBR))[f] =
        m.face_neighbor[(size_t)cel[s] * (D + 1) + f];
  }
  wave_lds_fence();
  for (int w = rank; w < nsl * (D + 1); w += nact) {
    const int s = w / (D + 1), f = w % (D + 1);
    const uint32_t nb = ((const uint32_t *)(base + s * C::BYTES + C::NBR))[f];
    ((int *)(base + s * C::BYTES + C::NBS))[f] =
        nb == INVALID ? -1 : cache_lookup<D, N, T>(cel, nsl, nb);
  }
  wave_lds_fence();
}

// inverse_map / in_unit_simplex on a cached inverse map
template <int D>
__device__ __forceinline__ bool inside_cached(const char *slot_inv, const double *x, double *r) {
  const double *a = (const double *)slot_inv;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double t = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) t += a[i * D + j] * x[j];
    r[i] = t + a[D * D + i];
  }
  return in_unit_simplex<D>(r);
}

// The head of relocate_wave with x known outside (particle_state.hpp:160-190: the face whose
// normal best matches the momentum, the neighbour across it) on the cached
// data of the lane's slot (slot < 0: from the mesh arrays).  Returns 1 when
// the neighbour contains x (elem, slot and r updated), 0 on the reference's
// "no best normal" error (err set), 2 when the cascade continues with the
// closest-vertex fan.
template <int D, int N, class T>
__device__ __forceinline__ int relocate_head(const DevMesh &m, const char *base,
                                             const uint32_t *cel, int nsl, const double *x,
                                             const double *p, uint32_t &elem, int &slot,
                                             FindCounters &cnt, int &err, double *r) {
  using C = StepCache<D, N, T>;
  double nrm[(D + 1) * D];
  const char *sb = base + (slot < 0 ? 0 : slot) * C::BYTES;
  if (slot >= 0) {
#pragma unroll
    for (int k = 0; k < (D + 1) * D; ++k) nrm[k] = ((const double *)(sb + C::NRM))[k];
  } else {
#pragma unroll
    for (int k = 0; k < (D + 1) * D; ++k) nrm[k] = m.face_normal[(size_t)elem * (D + 1) * D + k];
  }
  int closest = -1;
  double max_ip = 0;
#pragma unroll
  for (int f = 0; f <= D; ++f) {
    double ip = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) ip += nrm[f * D + i] * p[i];
    if (ip > max_ip) {
      closest = f;
      max_ip = ip;
    }
  }
  if (closest == -1) {
    err = 1;
    return 0;
  }
  uint32_t possible;
  int ps = -1;
  if (slot >= 0) {
    possible = ((const uint32_t *)(sb + C::NBR))[closest];
    ps = ((const int *)(sb + C::NBS))[closest];
  } else {
    possible = m.face_neighbor[(size_t)elem * (D + 1) + closest];
  }
  if (possible != INVALID) {
    const bool in = ps >= 0 ? inside_cached<D>(base + ps * C::BYTES + C::INV, x, r)
                            : in_element<D>(m, possible, x, r);
    if (in) {
      cnt.neighbor++;
      elem = possible;
      slot = ps >= 0 ? ps : cache_lookup<D, N, T>(cel, nsl, possible);
      return 1;
    }
  }
  return 2;
}

// One stage of the fused step on the wave's cached elements (k_lsrk4_step_c,
// k_lsrk4_ckpt): the update, the containment test and the cascade.  Returns
// false if the particle left the mesh (slot is then -1).
template <int D, int N, class T, bool MF>
__device__ __forceinline__ bool cached_stage(const DevMesh &m, const T *__restrict__ coef,
                                             char *base, const uint32_t *cel, int nsl, int s,
                                             double dt, double c, double qi, double mi,
                                             double *xi, double *pi, double *rx, double *rp,
                                             double *r, uint32_t &e, int &slot,
                                             FindCounters &cnt, int &e_flag, unsigned &n_mf) {
  using C = StepCache<D, N, T>;
  // coefficients: scalar loads for an element-uniform wave, LDS when every
  // active lane's element is cached, else per-lane loads
  const bool cached = __ballot(slot < 0) == 0;
  double v[D], F[D];
  velocity<D>(pi, mi, c, v);
  const uint32_t e_uni = wave_uniform_element(e);
  const T *cc = (const T *)(base + (slot < 0 ? 0 : slot) * C::BYTES);
  if constexpr (sizeof(T) == 8) {
    double f6[6], el[3], mag[3];
    if (e_uni != INVALID)
      interp6<D, N, double>(as_constant(coef),
                            (uint32_t)__builtin_amdgcn_readfirstlane((int)e_uni), r, f6);
    else if (cached)
      interp6<D, N, double>(cc, 0u, r, f6);
    else
      interp6<D, N, double>(coef, e, r, f6);
    lorentz<D, double>(qi, f6, v, F, el, mag);
  } else {
    float rf[D], f6[6], vf[D], Ff[D], el[3], mag[3];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      rf[k] = (float)r[k];
      vf[k] = (float)v[k];
    }
    if (MF && cached && __ballot(1) == ~0ull) {
      // MF: full waves with every element cached contract on the matrix cores
      interp6_mfma_slots<D, N>(base, slot, rf, f6);
      ++n_mf;
    } else if (e_uni != INVALID)
      interp6<D, N, float>(as_constant(coef),
                           (uint32_t)__builtin_amdgcn_readfirstlane((int)e_uni), rf, f6);
    else if (cached)
      interp6<D, N, float>(cc, 0u, rf, f6);
    else
      interp6<D, N, float>(coef, e, rf, f6);
    lorentz<D, float>((float)qi, f6, vf, Ff, el, mag);
#pragma unroll
    for (int k = 0; k < D; ++k) F[k] = (double)Ff[k];
  }
  const double a = c_rk4a[s], b = c_rk4b[s];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    rx[k] = a * rx[k] + dt * v[k];
    rp[k] = a * rp[k] + dt * F[k];
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xi[k] += b * rx[k];
    pi[k] += b * rp[k];
  }
  // the first test of find_new_containing_element, then the rest of the
  // cascade
  bool in;
  if (e_uni != INVALID)
    in = still_inside<D>(m, e, e_uni, xi, r);
  else if (slot >= 0)
    in = inside_cached<D>(base + slot * C::BYTES + C::INV, xi, r);
  else
    in = in_element<D>(m, e, xi, r);
  bool ok = true;
  int head = 1;
  if (in)
    cnt.same++;
  else
    head = relocate_head<D, N, T>(m, base, cel, nsl, xi, pi, e, slot, cnt, e_flag, r);
  // the rest of the cascade, one particle at a time with the whole wave
  wave_cascade_tail<D>(m, xi, e, cnt, r, head);
  if (head == 0) {
    ok = false;
  } else if (head == 2) {
    // no element contains x: periodic trip or death (boundary_hit)
    ok = boundary_hit<D>(m, xi, pi, e, cnt, e_flag, r);
    slot = ok ? cache_lookup<D, N, T>(cel, nsl, e) : -1;
  } else if (head == 3) {
    slot = cache_lookup<D, N, T>(cel, nsl, e);   // found by the fan or the grid
  }
  if (!ok) slot = -1;
  return ok;
}

// Slots for the elements the wave's live lanes entered, in slots no lane holds.
template <int D, int N, class T>
__device__ __forceinline__ void cache_refresh(const DevMesh &m, const T *__restrict__ coef,
                                              char *base, uint32_t *cel, int &nsl, bool live,
                                              uint32_t e, int &slot) {
  const uint64_t need = __ballot(live && slot < 0);
  if (!need) return;
  uint32_t used = 0;
  for (int j = 0; j < nsl; ++j)
    if (__ballot(slot == j)) used |= 1u << j;
  cache_update<D, N, T>(m, coef, base, cel, nsl, used, need, e, slot);
}

template <int D, int N, class T, bool MF>
__global__ __launch_bounds__(256, (step_min_waves<N, T>())) void k_lsrk4_step_c(
    DevMesh m, const T *__restrict__ coef, int64_t n, double dt, double c,
    double *__restrict__ x, double *__restrict__ p, const double *__restrict__ q,
    const double *__restrict__ mass, uint32_t *__restrict__ elem,
    uint32_t *__restrict__ death, unsigned long long *counters, int *err,
    unsigned int *n_dead, const uint32_t *__restrict__ order, unsigned long long *mf_waves) {
  using C = StepCache<D, N, T>;
  __shared__ uint4 lds[4][C::SLOTS * C::BYTES / 16];
  __shared__ uint32_t lds_el[4][C::SLOTS];
  const int wv = threadIdx.x >> 6;
  char *base = (char *)&lds[wv][0];
  uint32_t *cel = lds_el[wv];
  // thread t steps particle order[t] (t with a null order)
  const int64_t t =
      (order ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x) * blockDim.x + threadIdx.x;
  FindCounters cnt;
  unsigned n_mf = 0;   // stages this particle was interpolated on the matrix cores
  const bool live = t < n;
  const int64_t i = live ? (order ? (int64_t)order[t] : t) : 0;
  uint32_t e = live ? elem[i] : INVALID;
  const bool located = e < (uint32_t)m.K;
  // slots: the wave's start elements (lowest lane's first), then their face
  // neighbours, as many as fit
  int slot = -1, nsl = 0;
  cache_update<D, N, T>(m, coef, base, cel, nsl, 0u, __ballot(located), e, slot);
  if (live) {
    double xi[D], pi[D], r[D], rx[D], rp[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xi[k] = x[i * D + k];
      pi[k] = p[i * D + k];
      rx[k] = 0;
      rp[k] = 0;
    }
    const double qi = q[i], mi = mass[i];
    uint32_t died = 0;
    int e_flag = 0;
    if (!located) died = 1;   // not located: removed at the first stage
    else if (slot >= 0) inside_cached<D>(base + slot * C::BYTES + C::INV, xi, r);
    else inverse_map<D>(m, e, xi, r);
#pragma unroll 1
    for (int s = 0; s < 5 && !died; ++s) {
      if (!cached_stage<D, N, T, MF>(m, coef, base, cel, nsl, s, dt, c, qi, mi, xi, pi, rx, rp,
                                     r, e, slot, cnt, e_flag, n_mf))
        died = (uint32_t)s + 1;
      if (s < 4) cache_refresh<D, N, T>(m, coef, base, cel, nsl, !died, e, slot);
    }
    if (e_flag) atomicOr(err, 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[i * D + k] = xi[k];
      p[i * D + k] = pi[k];
    }
    elem[i] = e;
    death[i] = died;
    if (died) atomicAdd(n_dead, 1u);
  }
  if (MF && mf_waves) {
    unsigned long long v = n_mf;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(mf_waves, v);
  }
  flush_counters(cnt, counters);
}

// ---------------------------------------------------------------------------
// adjoint of the push
// ---------------------------------------------------------------------------
// Forward without compaction, nsteps in one launch (particles are independent
// in static fields).  A particle that leaves the mesh keeps its slot: death[i]
// = 5*step + stage (1-based) and elem[i] = INVALID.  Every ckpt_every steps the
// step-start state (x, p, element) is written to the checkpoint arrays,
// segment-major: ckpt_xp[(seg*n + i)*2D + (0..D-1: x, D..2D-1: p)].
template <int D, int N, class T>
__global__ __launch_bounds__(256) void k_lsrk4_ckpt(
    DevMesh m, const T *__restrict__ coef, int64_t n, int nsteps, int ckpt_every, double dt,
    double c, double *__restrict__ x, double *__restrict__ p, const double *__restrict__ q,
    const double *__restrict__ mass, uint32_t *__restrict__ elem, uint32_t *__restrict__ death,
    double *__restrict__ ckpt_xp, uint32_t *__restrict__ ckpt_e, unsigned long long *counters,
    int *err, const uint32_t *__restrict__ order) {
  // the stages run on the wave's LDS element cache (k_lsrk4_step_c), kept
  // across the steps: thread t steps particle order[t] (the element order at
  // the start), and the wave's particles travel together
  using C = StepCache<D, N, T>;
  __shared__ uint4 lds[4][C::SLOTS * C::BYTES / 16];
  __shared__ uint32_t lds_el[4][C::SLOTS];
  const int wv = threadIdx.x >> 6;
  char *base = (char *)&lds[wv][0];
  uint32_t *cel = lds_el[wv];
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  FindCounters cnt;
  unsigned n_mf = 0;
  const bool live = t < n;
  const int64_t i = live ? (int64_t)order[t] : 0;
  uint32_t e = live ? elem[i] : INVALID;
  const bool located = e < (uint32_t)m.K;
  int slot = -1, nsl = 0;
  cache_update<D, N, T>(m, coef, base, cel, nsl, 0u, __ballot(located), e, slot);
  if (live) {
    double xi[D], pi[D], r[D], rx[D], rp[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xi[k] = x[i * D + k];
      pi[k] = p[i * D + k];
    }
    const double qi = q[i], mi = mass[i];
    uint32_t died = 0;
    int e_flag = 0;
    if (!located) died = 1;
    else if (slot >= 0) inside_cached<D>(base + slot * C::BYTES + C::INV, xi, r);
    else inverse_map<D>(m, e, xi, r);
#pragma unroll 1
    for (int step = 0; step < nsteps && !died; ++step) {
      if (ckpt_xp && step % ckpt_every == 0) {
        const size_t seg = (size_t)(step / ckpt_every);
        double *dst = ckpt_xp + (seg * (size_t)n + (size_t)i) * 2 * D;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          dst[k] = xi[k];
          dst[D + k] = pi[k];
        }
        ckpt_e[seg * (size_t)n + (size_t)i] = e;
      }
#pragma unroll
      for (int k = 0; k < D; ++k) {
        rx[k] = 0;
        rp[k] = 0;
      }
#pragma unroll 1
      for (int s = 0; s < 5 && !died; ++s) {
        if (!cached_stage<D, N, T, false>(m, coef, base, cel, nsl, s, dt, c, qi, mi, xi, pi, rx,
                                          rp, r, e, slot, cnt, e_flag, n_mf))
          died = (uint32_t)(5 * step + s + 1);
        cache_refresh<D, N, T>(m, coef, base, cel, nsl, !died, e, slot);
      }
    }
    if (e_flag) atomicOr(err, 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[i * D + k] = xi[k];
      p[i * D + k] = pi[k];
    }
    elem[i] = died ? INVALID : e;
    death[i] = died;
  }
  flush_counters(cnt, counters);
}

// Adds J_f(y)^T kb to yb for the particle right-hand side f(x, p) = [v(p),
// q E(x) + v(p) x q B(x)] at the stage input y = (x, p) in element e:
//   vb = kb_x + (q B) x kb_p,   Eb = q kb_p,   Bb = q (kb_p x v),
//   xb += A^T grad_r (Eb . E + Bb . B),        A = d r / d x (inverse map),
//   pb += (c/s) (vb - p (p . vb) / s^2),       s = sqrt(m^2 c^2 + |p|^2).
// With T = float the field values and gradients are evaluated in FP32 on a
// power-of-two rescaled weight vector; the accumulation stays FP64.
// cf: the element's 6 x coef_stride field coefficients, am: its inverse map
// (D x D + D), both in global memory or in the wave's LDS slots.
template <int D, int N, class T, class CP, class AP>
__device__ __forceinline__ void rhs_adjoint(CP cf, AP am, double c, double qi, double mi,
                                            const double *xi, const double *pi,
                                            const double *kbx, const double *kbp, double *xb,
                                            double *pb) {
  constexpr Basis<D, N> B{};
  constexpr int NP = Basis<D, N>::NP;
  double r[D], v[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double t = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) t += am[i * D + j] * xi[j];
    r[i] = t + am[D * D + i];
  }
  velocity<D>(pi, mi, c, v);
  double v3[3], k3[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v3[k] = k < D ? v[k] : 0.0;
    k3[k] = k < D ? kbp[k] : 0.0;
  }
  double w6[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) w6[k] = qi * k3[k];
  w6[3] = qi * (k3[1] * v3[2] - k3[2] * v3[1]);
  w6[4] = qi * (k3[2] * v3[0] - k3[0] * v3[2]);
  w6[5] = qi * (k3[0] * v3[1] - k3[1] * v3[0]);
  double wmax = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) wmax = fmax(wmax, fabs(w6[k]));
  int wexp = 0;
  if (wmax > 0) frexp(wmax, &wexp);
  T ws[6], rt[D];
#pragma unroll
  for (int k = 0; k < 6; ++k) ws[k] = (T)ldexp(w6[k], -wexp);
#pragma unroll
  for (int k = 0; k < D; ++k) rt[k] = (T)r[k];
  T pw[D][N + 1];
#pragma unroll
  for (int ii = 0; ii < D; ++ii) {
    pw[ii][0] = T(1);
#pragma unroll
    for (int k = 1; k <= N; ++k) pw[ii][k] = pw[ii][k - 1] * rt[ii];
  }
  constexpr int NS = coef_stride<T>(NP);
  T bf[3] = {T(0), T(0), T(0)}, g[D];
#pragma unroll
  for (int k = 0; k < D; ++k) g[k] = T(0);
  // the coefficients V monomials at a time, one 16-byte load per field row
  // when the rows allow it (same products and order as one at a time)
  constexpr int V = (NS * sizeof(T)) % 16 == 0 ? 16 / (int)sizeof(T) : 1;
  struct alignas(V * sizeof(T)) Vec {
    T v[V];
  };
#pragma unroll
  for (int k0 = 0; k0 < NP; k0 += V) {
    Vec cv[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) cv[f] = *reinterpret_cast<const Vec *>(&cf[f * NS + k0]);
#pragma unroll
  for (int kk = 0; kk < V && k0 + kk < NP; ++kk) {
    const int k = k0 + kk;
    T phi = pw[0][B.e[k][0]];
#pragma unroll
    for (int ii = 1; ii < D; ++ii) phi = phi * pw[ii][B.e[k][ii]];
#pragma unroll
    for (int f = 0; f < 3; ++f) bf[f] = fma(cv[3 + f].v[kk], phi, bf[f]);
    T wk = T(0);
#pragma unroll
    for (int f = 0; f < 6; ++f) wk = fma(ws[f], cv[f].v[kk], wk);
#pragma unroll
    for (int jj = 0; jj < D; ++jj) {
      const int ej = B.e[k][jj];
      if (ej > 0) {
        T d = T(ej) * pw[jj][ej - 1];
#pragma unroll
        for (int ii = 0; ii < D; ++ii)
          if (ii != jj) d = d * pw[ii][B.e[k][ii]];
        g[jj] = fma(wk, d, g[jj]);
      }
    }
  }
  }
  // x: A^T grad_r, A row-major in the inverse map
#pragma unroll
  for (int jj = 0; jj < D; ++jj) {
    double acc = 0;
#pragma unroll
    for (int ii = 0; ii < D; ++ii) acc += (double)g[ii] * am[ii * D + jj];
    xb[jj] += ldexp(acc, wexp);
  }
  // v: kb_x + (q B) x kb_p, truncated to D components
  double qb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) qb[k] = qi * (double)bf[k];
  double vb3[3];
  vb3[0] = qb[1] * k3[2] - qb[2] * k3[1];
  vb3[1] = qb[2] * k3[0] - qb[0] * k3[2];
  vb3[2] = qb[0] * k3[1] - qb[1] * k3[0];
  double vb[D], pv = 0, pp = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    vb[k] = kbx[k] + vb3[k];
    pv += pi[k] * vb[k];
    pp += pi[k] * pi[k];
  }
  const double s2 = mi * mi * c * c + pp, s = sqrt(s2), f = c / s;
#pragma unroll
  for (int k = 0; k < D; ++k) pb[k] += f * (vb[k] - pi[k] * (pv / s2));
}

// Reverse pass over all steps for every live particle, one thread per
// particle.  For each checkpoint segment, last first: recompute the segment's
// step-start states into ybuf, then for each step, last first, recompute its
// five stage inputs into sbuf and run the stages backwards:
//   rb += b_s yb;  kb = dt rb;  rb *= a_s;  yb += J_f(y_{s-1})^T kb.
// Particles that died get zero gradients (they are not in the final state).
// FP64: capped at 256 VGPRs (2 waves per SIMD instead of 1 at 282); a few
// spills, 28% faster on the cylinder adjoint (1140 -> 822 ms for 200 steps).
template <int D, int N, class T>
__global__ __launch_bounds__(256, (sizeof(T) == 8 ? 2 : 1)) void k_lsrk4_adjoint(
    DevMesh m, const T *__restrict__ coef, int64_t n, int nsteps, int ckpt_every, double dt,
    double c, const double *__restrict__ q, const double *__restrict__ mass,
    const uint32_t *__restrict__ death, const double *__restrict__ ckpt_xp,
    const uint32_t *__restrict__ ckpt_e, double *__restrict__ ybuf, uint32_t *__restrict__ ebuf,
    double *__restrict__ sbuf, uint32_t *__restrict__ sebuf, double *__restrict__ x_bar,
    double *__restrict__ p_bar, int *err, const uint32_t *__restrict__ order) {
  // the recomputed stages and the adjoint right-hand sides read the elements
  // from the wave's LDS slots (k_lsrk4_step_c), kept across the segments
  using C = StepCache<D, N, T>;
  __shared__ uint4 lds[4][C::SLOTS * C::BYTES / 16];
  __shared__ uint32_t lds_el[4][C::SLOTS];
  const int wv = threadIdx.x >> 6;
  char *base = (char *)&lds[wv][0];
  uint32_t *cel = lds_el[wv];
  int nsl = 0;
  // thread t takes particle order[t] (the element order at the start), so a
  // wave's particles share elements
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = (int64_t)order[t];
  if (death[i]) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x_bar[i * D + k] = 0;
      p_bar[i * D + k] = 0;
    }
    return;
  }
  double ax[D], ap[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    ax[k] = x_bar[i * D + k];
    ap[k] = p_bar[i * D + k];
  }
  const double qi = q[i], mi = mass[i];
  FindCounters cnt;
  unsigned n_mf = 0;
  int e_flag = 0, lost = 0;
  // the lane's slot of element e (wave-collective: the cached copy, else one
  // inserted now; -1 when the slots are full)
  auto slot_of = [&](uint32_t e) {
    int sl = cache_lookup<D, N, T>(cel, nsl, e);
    cache_refresh<D, N, T>(m, coef, base, cel, nsl, true, e, sl);
    return sl;
  };
  // unit coordinates of x in element e through its slot
  auto coords = [&](uint32_t e, int sl, const double *xi, double *r) {
    if (sl >= 0) inside_cached<D>(base + sl * C::BYTES + C::INV, xi, r);
    else inverse_map<D>(m, e, xi, r);
  };
  // checkpoints are per particle (i); the recompute buffers per thread (t)
  const size_t nn = (size_t)n, ii = (size_t)i, tt = (size_t)t;
  const int nseg = (nsteps + ckpt_every - 1) / ckpt_every;
#pragma unroll 1
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int steps = min(ckpt_every, nsteps - seg * ckpt_every);
    double xi[D], pi[D], r[D], rx[D], rp[D];
    const double *src = ckpt_xp + ((size_t)seg * nn + ii) * 2 * D;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xi[k] = src[k];
      pi[k] = src[D + k];
    }
    uint32_t e = ckpt_e[(size_t)seg * nn + ii];
    int slot = slot_of(e);
    coords(e, slot, xi, r);
#pragma unroll 1
    for (int k = 0; k < steps; ++k) {
      double *dst = ybuf + ((size_t)k * nn + tt) * 2 * D;
#pragma unroll
      for (int kk = 0; kk < D; ++kk) {
        dst[kk] = xi[kk];
        dst[D + kk] = pi[kk];
      }
      ebuf[(size_t)k * nn + tt] = e;
      if (k + 1 == steps) break;
#pragma unroll
      for (int kk = 0; kk < D; ++kk) {
        rx[kk] = 0;
        rp[kk] = 0;
      }
#pragma unroll 1
      for (int s = 0; s < 5; ++s) {
        if (!cached_stage<D, N, T, false>(m, coef, base, cel, nsl, s, dt, c, qi, mi, xi, pi, rx,
                                          rp, r, e, slot, cnt, e_flag, n_mf))
          lost = 1;
        cache_refresh<D, N, T>(m, coef, base, cel, nsl, true, e, slot);
      }
    }
#pragma unroll 1
    for (int k = steps - 1; k >= 0; --k) {
      const double *y = ybuf + ((size_t)k * nn + tt) * 2 * D;
#pragma unroll
      for (int kk = 0; kk < D; ++kk) {
        xi[kk] = y[kk];
        pi[kk] = y[D + kk];
        rx[kk] = 0;
        rp[kk] = 0;
      }
      e = ebuf[(size_t)k * nn + tt];
      slot = slot_of(e);
      coords(e, slot, xi, r);
#pragma unroll 1
      for (int s = 0; s < 5; ++s) {
        double *st = sbuf + ((size_t)s * nn + tt) * 2 * D;
#pragma unroll
        for (int kk = 0; kk < D; ++kk) {
          st[kk] = xi[kk];
          st[D + kk] = pi[kk];
        }
        sebuf[(size_t)s * nn + tt] = e;
        if (s == 4) break;   // the last stage's output is not needed
        if (!cached_stage<D, N, T, false>(m, coef, base, cel, nsl, s, dt, c, qi, mi, xi, pi, rx,
                                          rp, r, e, slot, cnt, e_flag, n_mf))
          lost = 1;
        cache_refresh<D, N, T>(m, coef, base, cel, nsl, true, e, slot);
      }
      double rbx[D], rbp[D];
#pragma unroll
      for (int kk = 0; kk < D; ++kk) {
        rbx[kk] = 0;
        rbp[kk] = 0;
      }
#pragma unroll 1
      for (int s = 4; s >= 0; --s) {
        const double a = c_rk4a[s], b = c_rk4b[s];
        double kbx[D], kbp[D];
#pragma unroll
        for (int kk = 0; kk < D; ++kk) {
          rbx[kk] += b * ax[kk];
          rbp[kk] += b * ap[kk];
          kbx[kk] = dt * rbx[kk];
          kbp[kk] = dt * rbp[kk];
          rbx[kk] *= a;
          rbp[kk] *= a;
        }
        const double *st = sbuf + ((size_t)s * nn + tt) * 2 * D;
        double xs[D], ps[D];
#pragma unroll
        for (int kk = 0; kk < D; ++kk) {
          xs[kk] = st[kk];
          ps[kk] = st[D + kk];
        }
        // FP64: from the slots, in rounds when the wave's elements outnumber
        // them (each round frees the slots of the lanes it served); FP32C:
        // from the mesh arrays (its FP32 rows read from LDS measured slower,
        // 129 -> 148 ms for the 4M-particle 200-step adjoint, against
        // 239 -> 179 ms for FP64)
        const uint32_t es = sebuf[(size_t)s * nn + tt];
        if constexpr (sizeof(T) == 4) {
          rhs_adjoint<D, N, T>(coef + (size_t)es * 6 * coef_stride<T>(Basis<D, N>::NP),
                               m.inv_map + (size_t)es * (D * D + D), c, qi, mi, xs, ps, kbx,
                               kbp, ax, ap);
          continue;
        }
        bool done = false;
        do {
          int ss = done ? -1 : cache_lookup<D, N, T>(cel, nsl, es);
          cache_refresh<D, N, T>(m, coef, base, cel, nsl, !done, es, ss);
          if (!done && ss >= 0) {
            const char *sb = base + ss * C::BYTES;
            rhs_adjoint<D, N, T>((const T *)sb, (const double *)(sb + C::INV), c, qi, mi, xs,
                                 ps, kbx, kbp, ax, ap);
            done = true;
          }
        } while (__ballot(!done));
      }
    }
  }
  if (e_flag || lost) atomicOr(err, 2);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    x_bar[i * D + k] = ax[k];
    p_bar[i * D + k] = ap[k];
  }
}

// p_bar = (dv/dp)^T v_bar for v = c p / sqrt(m^2 c^2 + |p|^2)
template <int VD>
__global__ void k_velocities_adjoint(int64_t n, double c, const double *__restrict__ p,
                                     const double *__restrict__ mass,
                                     const double *__restrict__ v_bar, double *__restrict__ p_bar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double pi[VD], vb[VD], pv = 0, pp = 0;
#pragma unroll
  for (int k = 0; k < VD; ++k) {
    pi[k] = p[i * VD + k];
    vb[k] = v_bar[i * VD + k];
    pv += pi[k] * vb[k];
    pp += pi[k] * pi[k];
  }
  const double mi = mass[i];
  const double s2 = mi * mi * c * c + pp, s = sqrt(s2), f = c / s;
#pragma unroll
  for (int k = 0; k < VD; ++k) p_bar[i * VD + k] = f * (vb[k] - pi[k] * (pv / s2));
}

__global__ void k_stage_flags(int64_t n, const uint32_t *__restrict__ death, uint32_t s,
                              uint8_t *__restrict__ flags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = death[i] == s ? 1 : 0;
}

// One LSRK4 stage of the averaging pusher's step for the particles still
// alive (death == 0), in the per-stage path's arithmetic: residual = a
// residual + dt k, y = y + b residual (timestep.py, cloud.py:554-603), then
// relocalization; a particle that leaves is frozen with death = stage + 1 and
// elem = INVALID.
template <int D>
__global__ __launch_bounds__(256) void k_avg_stage(DevMesh m, int64_t n, int s, double dt,
                                                   double *__restrict__ x, double *__restrict__ p,
                                                   double *__restrict__ rx, double *__restrict__ rp,
                                                   const double *__restrict__ v,
                                                   const double *__restrict__ F,
                                                   uint32_t *__restrict__ elem,
                                                   uint32_t *__restrict__ death,
                                                   unsigned long long *counters, int *err,
                                                   unsigned int *n_dead) {
  FindCounters cnt{};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // every lane takes part in the wave's fan and grid searches
  const bool live = i < n && death[i] == 0;
  double xi[D] = {}, pi[D] = {}, r[D];
  uint32_t e = INVALID;
  if (live) {
    const double a = c_rk4a[s], b = c_rk4b[s];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double rxk = a * rx[i * D + k] + dt * v[i * D + k];
      const double rpk = a * rp[i * D + k] + dt * F[i * D + k];
      rx[i * D + k] = rxk;
      rp[i * D + k] = rpk;
      xi[k] = x[i * D + k] + b * rxk;
      pi[k] = p[i * D + k] + b * rpk;
    }
    e = elem[i];
  }
  int e_flag = 0;
  const bool alive = relocate_wave<D>(m, live, false, xi, pi, e, cnt, e_flag, r);
  if (live) {
    if (e_flag) atomicOr(err, 1);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[i * D + k] = xi[k];
      p[i * D + k] = pi[k];
    }
    elem[i] = alive ? e : INVALID;
    if (!alive) {
      death[i] = (uint32_t)s + 1;
      atomicAdd(n_dead, 1u);
    }
  }
  flush_counters(cnt, counters);
}

static inline unsigned blocks_for(int64_t n, int bs = 256) {
  return (unsigned)((n + bs - 1) / bs > 0 ? (n + bs - 1) / bs : 1);
}

static int order_of(const pyr_mesh *h) {
  // the exponents must be the graded node_tuples() order for some N
  const int D = h->dim;
  for (int N = 1; N <= 6; ++N) {
    int np = D == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
    if (np != h->Np) continue;
    std::vector<int32_t> ex;
    for (int total = 0; total <= N; ++total) {
      if (D == 2)
        for (int j = 0; j <= total; ++j) {
          ex.push_back(total - j);
          ex.push_back(j);
        }
      else
        for (int j = 0; j <= total; ++j)
          for (int l = 0; l <= total - j; ++l) {
            ex.push_back(total - j - l);
            ex.push_back(j);
            ex.push_back(l);
          }
    }
    if (ex == h->h_exps) return N;
  }
  throw Error(PYR_ERR_INVALID_ARG,
              "unsupported monomial basis (need graded node tuples, order 1..6 in 2D, "
              "1..4 in 3D)");
}

template <class F>
static void dispatch_order(int D, int N, F &&f) {
#define PYR_CASE(DD, NN) \
  if (D == DD && N == NN) return f(std::integral_constant<int, DD>{}, std::integral_constant<int, NN>{});
  PYR_CASE(2, 1) PYR_CASE(2, 2) PYR_CASE(2, 3) PYR_CASE(2, 4) PYR_CASE(2, 5) PYR_CASE(2, 6)
  PYR_CASE(3, 1) PYR_CASE(3, 2) PYR_CASE(3, 3) PYR_CASE(3, 4)
#undef PYR_CASE
  throw Error(PYR_ERR_INVALID_ARG, "unsupported dimension/order");
}

// occ/stamp: restrict the FP64 transform to the marked elements (null: all)
static void modal_transform(pyr_mesh *h, const double *const fields[6], double *out64,
                            float *out32, const uint32_t *occ = nullptr, uint32_t stamp = 0) {
  Fields6 fl;
  for (int f = 0; f < 6; ++f) fl.f[f] = fields ? fields[f] : nullptr;
  const int64_t work = (int64_t)h->K * h->Np;
  if (out64) {
    hipLaunchKernelGGL(k_modal<double>, dim3(blocks_for(work)), dim3(256), 0, stream(), h->K,
                       h->Np, fl, h->minv.as<double>(), out64, occ, stamp);
    PYR_CHECK_LAUNCH();
  }
  if (out32) {
    const int64_t work32 = (int64_t)h->K * coef_stride<float>(h->Np);
    hipLaunchKernelGGL(k_modal<float>, dim3(blocks_for(work32)), dim3(256), 0, stream(), h->K,
                       h->Np, fl, h->minv.as<double>(), out32, (const uint32_t *)nullptr, 0u);
    PYR_CHECK_LAUNCH();
  }
}

// occ[elem[i]] = stamp for the located particles
__global__ void k_mark_occupied(int64_t n, const uint32_t *__restrict__ elem, uint32_t K,
                                uint32_t *__restrict__ occ, uint32_t stamp) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t e = elem[i];
  if (e < K) occ[e] = stamp;
}

}  // namespace pyr

using namespace pyr;

extern "C" {

int pyr_forces(pyr_mesh *mesh, int64_t n, int vdim, const double *const fields[6],
               const double *x, const uint32_t *elem, const double *q, const double *v,
               double *force, double *vis_e, double *vis_b, double *vis_el, double *vis_mag) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n >= 0, "invalid argument");
    PYR_REQUIRE(vdim == mesh->dim, "vdim must equal the mesh dimension");
    if (n == 0) return;
    PYR_REQUIRE(x && elem && q && v && force, "null pointer");
    const bool vis = vis_e != nullptr;
    PYR_REQUIRE(!vis || (vis_b && vis_el && vis_mag), "all four vis arrays or none");
    const int N = order_of(mesh);
    mesh->s_modal.reserve(sizeof(double) * (size_t)mesh->K * 6 * mesh->Np);
    // modal coefficients of the occupied elements only: the forces read no
    // other row, and a beam occupies a small share of the mesh.  A stamp that
    // changes per call marks them, so the marks need no clearing.
    const size_t occ_bytes = sizeof(uint32_t) * (size_t)mesh->K;
    if (mesh->s_occ.bytes < occ_bytes || mesh->occ_stamp == UINT32_MAX) {
      mesh->s_occ.reserve(occ_bytes);
      PYR_HIP(hipMemsetAsync(mesh->s_occ.ptr, 0, occ_bytes, stream()));
      mesh->occ_stamp = 0;
    }
    const uint32_t stamp = ++mesh->occ_stamp;
    hipLaunchKernelGGL(k_mark_occupied, dim3(blocks_for(n)), dim3(256), 0, stream(), n, elem,
                       (uint32_t)mesh->K, mesh->s_occ.as<uint32_t>(), stamp);
    PYR_CHECK_LAUNCH();
    modal_transform(mesh, fields, mesh->s_modal.as<double>(), nullptr,
                    mesh->s_occ.as<uint32_t>(), stamp);
    dispatch_order(mesh->dim, N, [&](auto DD, auto NN) {
      constexpr int D_ = decltype(DD)::value, N_ = decltype(NN)::value;
      hipLaunchKernelGGL((k_forces<D_, N_>), dim3(blocks_for(n)), dim3(256), 0, stream(),
                         mesh->dev, mesh->s_modal.as<double>(), n, x, elem, q, v, force,
                         vis_e, vis_b, vis_el, vis_mag);
      PYR_CHECK_LAUNCH();
    });
  });
}

int pyr_set_static_fields(pyr_mesh *mesh, const double *const fields[6]) {
  return guarded([&] {
    PYR_REQUIRE(mesh, "null mesh");
    order_of(mesh);
    const size_t cnt = (size_t)mesh->K * 6 * mesh->Np;
    mesh->sf64.reserve(cnt * sizeof(double));
    mesh->sf32.reserve((size_t)mesh->K * 6 * coef_stride<float>(mesh->Np) * sizeof(float));
    modal_transform(mesh, fields, mesh->sf64.as<double>(), mesh->sf32.as<float>());
    const size_t nodes = (size_t)mesh->K * mesh->Np;
    mesh->sf_nodal.reserve(6 * nodes * sizeof(double));
    for (int f = 0; f < 6; ++f) {
      mesh->sf_has[f] = fields[f] != nullptr;
      if (fields[f])
        PYR_HIP(hipMemcpyAsync(mesh->sf_nodal.as<double>() + f * nodes, fields[f],
                               nodes * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    }
    mesh->sf_set = true;
  });
}

int pyr_lsrk4_steps(pyr_mesh *mesh, int precision, int nsteps, double dt, double vacuum_c,
                    int64_t *n_inout, double *x, double *p, double *q, double *mass,
                    uint32_t *elem, int64_t *ids, unsigned long long *counters) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n_inout && nsteps >= 0, "invalid argument");
    PYR_REQUIRE(precision == PYR_FP64 || precision == PYR_FP32C, "unknown precision");
    PYR_REQUIRE(mesh->sf_set, "pyr_set_static_fields has not been called for this mesh");
    int64_t n = *n_inout;
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    if (n == 0 || nsteps == 0) return;
    PYR_REQUIRE(x && p && q && mass && elem, "null pointer");
    const int D = mesh->dim, N = order_of(mesh);
    // scratch: death stage (u32), flags (u8), moves (2 x u32), err, n_dead
    mesh->s_flags.reserve((size_t)n * (sizeof(uint32_t) + 1));
    mesh->s_holes.reserve((size_t)n * 2 * sizeof(uint32_t));
    mesh->s_misc.reserve(64);
    uint32_t *death = mesh->s_flags.as<uint32_t>();
    uint8_t *flags = reinterpret_cast<uint8_t *>(death + n);
    uint32_t *mv_from = mesh->s_holes.as<uint32_t>();
    uint32_t *mv_to = mv_from + n;
    int *err = mesh->s_misc.as<int>();
    unsigned int *n_dead = reinterpret_cast<unsigned int *>(err + 1);
    unsigned long long *mf_waves = nullptr;
    if (mesh->push_mfma == 2) {
      mesh->mf_count.reserve(sizeof(unsigned long long));
      mf_waves = mesh->mf_count.as<unsigned long long>();
    }
    for (int step = 0; step < nsteps && n > 0; ++step) {
      PYR_HIP(hipMemsetAsync(err, 0, 2 * sizeof(int), stream()));
      // particles in the order of the last element sort of these n particles
      // (the deposition's), so that waves mostly share one element and read
      // its coefficients through scalar loads, with chunks of consecutive
      // blocks on one XCD (4M-particle cylinder: FP64 push 1.33 -> 1.11 ms,
      // FP32C 0.99 -> 0.95 ms).  A step with kills compacts the arrays and
      // changes n, which drops the order.
      const uint32_t *order = deposit_last_order(mesh, n);
      {
        ProfScope prof(PROF_PUSH);
        dispatch_order(D, N, [&](auto DD, auto NN) {
          constexpr int D_ = decltype(DD)::value, N_ = decltype(NN)::value;
          if (precision == PYR_FP64)
            hipLaunchKernelGGL((k_lsrk4_step_c<D_, N_, double, false>), dim3(blocks_for(n)),
                               dim3(256), 0, stream(), mesh->dev, mesh->sf64.as<double>(), n, dt,
                               vacuum_c, x, p, q, mass, elem, death, counters, err, n_dead, order,
                               nullptr);
          else if (mesh->push_mfma)
            hipLaunchKernelGGL((k_lsrk4_step_c<D_, N_, float, true>), dim3(blocks_for(n)),
                               dim3(256), 0, stream(), mesh->dev, mesh->sf32.as<float>(), n, dt,
                               vacuum_c, x, p, q, mass, elem, death, counters, err, n_dead, order,
                               mf_waves);
          else
            hipLaunchKernelGGL((k_lsrk4_step_c<D_, N_, float, false>), dim3(blocks_for(n)),
                               dim3(256), 0, stream(), mesh->dev, mesh->sf32.as<float>(), n, dt,
                               vacuum_c, x, p, q, mass, elem, death, counters, err, n_dead, order,
                               nullptr);
          PYR_CHECK_LAUNCH();
        });
      }
      int h[2] = {0, 0};
      read_step_flags(mesh, err, h);
      if (h[0]) throw Error(PYR_ERR_NO_BEST_NORMAL, "no best normal found--weird");
      if (h[1] == 0) continue;
      // the reference compacts after each stage (cloud.py:587-592); replay
      // the five compactions in stage order on the stage-of-death tags
      for (uint32_t s = 1; s <= 5 && n > 0; ++s) {
        hipLaunchKernelGGL(k_stage_flags, dim3(blocks_for(n)), dim3(256), 0, stream(), n,
                           death, s, flags);
        PYR_CHECK_LAUNCH();
        int64_t n_new = 0, n_moves = 0;
        int rc = pyr_compact_plan(n, flags, mv_from, mv_to, &n_new, &n_moves);
        if (rc) throw Error(rc, pyr_last_error());
        if (n_moves) {
          struct {
            void *buf;
            int64_t bytes;
          } arrays[] = {{x, 8 * D}, {p, 8 * D}, {q, 8}, {mass, 8}, {elem, 4},
                        {death, 4}, {ids, 8}};
          for (auto &a : arrays)
            if (a.buf) {
              rc = pyr_apply_moves(a.buf, a.bytes, n_moves, mv_from, mv_to);
              if (rc) throw Error(rc, pyr_last_error());
            }
        }
        n = n_new;
      }
    }
    *n_inout = n;
  });
}

int pyr_lsrk4_steps_average(pyr_mesh *mesh, int nsteps, double dt, double vacuum_c,
                            double radius, double alpha, double normalizer,
                            const double *integral_weights, int64_t *n_inout, double *x,
                            double *p, double *q, double *mass, uint32_t *elem, int64_t *ids,
                            unsigned long long *counters) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n_inout && nsteps >= 0, "invalid argument");
    PYR_REQUIRE(mesh->sf_set, "pyr_set_static_fields has not been called for this mesh");
    int64_t n = *n_inout;
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    if (n == 0 || nsteps == 0) return;
    PYR_REQUIRE(x && p && q && mass && elem && integral_weights, "null pointer");
    const int D = mesh->dim;
    const size_t nodes = (size_t)mesh->K * mesh->Np;
    const double *fields[6];
    for (int f = 0; f < 6; ++f)
      fields[f] = mesh->sf_has[f] ? mesh->sf_nodal.as<double>() + f * nodes : nullptr;
    // scratch: residuals rx, rp, velocities, forces (n x D each); death
    // stage, flags, moves, error flag, dead count
    const size_t nd = (size_t)n * D;
    mesh->s_avg.reserve(4 * nd * sizeof(double));
    double *rx = mesh->s_avg.as<double>(), *rp = rx + nd, *v = rp + nd, *F = v + nd;
    mesh->s_flags.reserve((size_t)n * (sizeof(uint32_t) + 1));
    mesh->s_holes.reserve((size_t)n * 2 * sizeof(uint32_t));
    mesh->s_misc.reserve(64);
    uint32_t *death = mesh->s_flags.as<uint32_t>();
    uint8_t *flags = reinterpret_cast<uint8_t *>(death + n);
    uint32_t *mv_from = mesh->s_holes.as<uint32_t>();
    uint32_t *mv_to = mv_from + n;
    int *err = mesh->s_misc.as<int>();
    unsigned int *n_dead = reinterpret_cast<unsigned int *>(err + 1);
    for (int step = 0; step < nsteps && n > 0; ++step) {
      // the residual starts as 0 * rhs (timestep.py); a_0 = 0 keeps it there
      PYR_HIP(hipMemsetAsync(rx, 0, 2 * (size_t)n * D * sizeof(double), stream()));
      PYR_HIP(hipMemsetAsync(death, 0, (size_t)n * sizeof(uint32_t), stream()));
      PYR_HIP(hipMemsetAsync(err, 0, 2 * sizeof(int), stream()));
      for (int s = 0; s < 5; ++s) {
        int rc = pyr_velocities(n, D, vacuum_c, p, mass, v);
        if (rc) throw Error(rc, pyr_last_error());
        average_forces_impl(mesh, radius, alpha, normalizer, n, D, fields, x, elem, q, v,
                            integral_weights, F, nullptr, nullptr);
        if (D == 2)
          hipLaunchKernelGGL(k_avg_stage<2>, dim3(blocks_for(n)), dim3(256), 0, stream(),
                             mesh->dev, n, s, dt, x, p, rx, rp, v, F, elem, death, counters, err,
                             n_dead);
        else
          hipLaunchKernelGGL(k_avg_stage<3>, dim3(blocks_for(n)), dim3(256), 0, stream(),
                             mesh->dev, n, s, dt, x, p, rx, rp, v, F, elem, death, counters, err,
                             n_dead);
        PYR_CHECK_LAUNCH();
      }
      int h[2] = {0, 0};
      read_step_flags(mesh, err, h);
      if (h[0]) throw Error(PYR_ERR_NO_BEST_NORMAL, "no best normal found--weird");
      if (h[1] == 0) continue;
      // compactions in stage order, as after each stage of the per-stage path
      for (uint32_t s = 1; s <= 5 && n > 0; ++s) {
        hipLaunchKernelGGL(k_stage_flags, dim3(blocks_for(n)), dim3(256), 0, stream(), n,
                           death, s, flags);
        PYR_CHECK_LAUNCH();
        int64_t n_new = 0, n_moves = 0;
        int rc = pyr_compact_plan(n, flags, mv_from, mv_to, &n_new, &n_moves);
        if (rc) throw Error(rc, pyr_last_error());
        if (n_moves) {
          struct {
            void *buf;
            int64_t bytes;
          } arrays[] = {{x, 8 * D}, {p, 8 * D}, {q, 8}, {mass, 8}, {elem, 4},
                        {death, 4}, {ids, 8}};
          for (auto &a : arrays)
            if (a.buf) {
              rc = pyr_apply_moves(a.buf, a.bytes, n_moves, mv_from, mv_to);
              if (rc) throw Error(rc, pyr_last_error());
            }
        }
        n = n_new;
      }
    }
    *n_inout = n;
  });
}

int pyr_lsrk4_forward(pyr_mesh *mesh, int precision, int nsteps, int ckpt_every, double dt,
                      double vacuum_c, int64_t n, double *x, double *p, const double *q,
                      const double *mass, uint32_t *elem, uint32_t *death, double *ckpt_xp,
                      uint32_t *ckpt_elem, unsigned long long *counters) {
  return guarded([&] {
    PYR_REQUIRE(mesh && nsteps >= 0 && ckpt_every > 0, "invalid argument");
    PYR_REQUIRE(precision == PYR_FP64 || precision == PYR_FP32C, "unknown precision");
    PYR_REQUIRE(mesh->sf_set, "pyr_set_static_fields has not been called for this mesh");
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    PYR_REQUIRE((ckpt_xp == nullptr) == (ckpt_elem == nullptr),
                "checkpoint positions/momenta and elements go together");
    if (n == 0) return;
    PYR_REQUIRE(x && p && q && mass && elem && death, "null pointer");
    const int D = mesh->dim, N = order_of(mesh);
    mesh->s_misc.reserve(64);
    int *err = mesh->s_misc.as<int>();
    PYR_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
    {
      ProfScope prof(PROF_PUSH);
      const uint32_t *order = element_order(mesh, n, elem);
      dispatch_order(D, N, [&](auto DD, auto NN) {
        constexpr int D_ = decltype(DD)::value, N_ = decltype(NN)::value;
        if (precision == PYR_FP64)
          hipLaunchKernelGGL((k_lsrk4_ckpt<D_, N_, double>), dim3(blocks_for(n)), dim3(256), 0,
                             stream(), mesh->dev, mesh->sf64.as<double>(), n, nsteps, ckpt_every,
                             dt, vacuum_c, x, p, q, mass, elem, death, ckpt_xp, ckpt_elem,
                             counters, err, order);
        else
          hipLaunchKernelGGL((k_lsrk4_ckpt<D_, N_, float>), dim3(blocks_for(n)), dim3(256), 0,
                             stream(), mesh->dev, mesh->sf32.as<float>(), n, nsteps, ckpt_every,
                             dt, vacuum_c, x, p, q, mass, elem, death, ckpt_xp, ckpt_elem,
                             counters, err, order);
        PYR_CHECK_LAUNCH();
      });
    }
    int h = 0;
    PYR_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, stream()));
    PYR_HIP(hipStreamSynchronize(stream()));
    if (h) throw Error(PYR_ERR_NO_BEST_NORMAL, "no best normal found--weird");
  });
}

int pyr_lsrk4_adjoint(pyr_mesh *mesh, int precision, int nsteps, int ckpt_every, double dt,
                      double vacuum_c, int64_t n, const double *q, const double *mass,
                      const uint32_t *death, const double *ckpt_xp, const uint32_t *ckpt_elem,
                      double *x_bar, double *p_bar) {
  return guarded([&] {
    PYR_REQUIRE(mesh && nsteps >= 0 && ckpt_every > 0, "invalid argument");
    PYR_REQUIRE(precision == PYR_FP64 || precision == PYR_FP32C, "unknown precision");
    PYR_REQUIRE(mesh->sf_set, "pyr_set_static_fields has not been called for this mesh");
    PYR_REQUIRE(n >= 0 && n < (int64_t)0x7FFFFFFF, "invalid particle count");
    if (n == 0 || nsteps == 0) return;
    PYR_REQUIRE(q && mass && death && ckpt_xp && ckpt_elem && x_bar && p_bar, "null pointer");
    const int D = mesh->dim, N = order_of(mesh);
    // ybuf: ckpt_every step-start states; sbuf: five stage inputs
    const size_t rec = (size_t)2 * D * sizeof(double) + sizeof(uint32_t);
    const size_t ybytes = (size_t)ckpt_every * (size_t)n * 2 * D * sizeof(double);
    const size_t ebytes = (size_t)ckpt_every * (size_t)n * sizeof(uint32_t);
    const size_t sbytes = (size_t)5 * (size_t)n * 2 * D * sizeof(double);
    const size_t sebytes = (size_t)5 * (size_t)n * sizeof(uint32_t);
    (void)rec;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    mesh->s_adj.reserve(al(ybytes) + al(ebytes) + al(sbytes) + al(sebytes));
    char *base = mesh->s_adj.as<char>();
    double *ybuf = reinterpret_cast<double *>(base);
    uint32_t *ebuf = reinterpret_cast<uint32_t *>(base + al(ybytes));
    double *sbuf = reinterpret_cast<double *>(base + al(ybytes) + al(ebytes));
    uint32_t *sebuf = reinterpret_cast<uint32_t *>(base + al(ybytes) + al(ebytes) + al(sbytes));
    mesh->s_misc.reserve(64);
    int *err = mesh->s_misc.as<int>();
    PYR_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
    {
      ProfScope prof(PROF_ADJOINT);
      // the element order at the start (first checkpoint): a wave's particles
      // share elements through the segments
      const uint32_t *order = element_order(mesh, n, ckpt_elem);
      dispatch_order(D, N, [&](auto DD, auto NN) {
        constexpr int D_ = decltype(DD)::value, N_ = decltype(NN)::value;
        if (precision == PYR_FP64)
          hipLaunchKernelGGL((k_lsrk4_adjoint<D_, N_, double>), dim3(blocks_for(n)), dim3(256),
                             0, stream(), mesh->dev, mesh->sf64.as<double>(), n, nsteps,
                             ckpt_every, dt, vacuum_c, q, mass, death, ckpt_xp, ckpt_elem, ybuf,
                             ebuf, sbuf, sebuf, x_bar, p_bar, err, order);
        else
          hipLaunchKernelGGL((k_lsrk4_adjoint<D_, N_, float>), dim3(blocks_for(n)), dim3(256),
                             0, stream(), mesh->dev, mesh->sf32.as<float>(), n, nsteps,
                             ckpt_every, dt, vacuum_c, q, mass, death, ckpt_xp, ckpt_elem, ybuf,
                             ebuf, sbuf, sebuf, x_bar, p_bar, err, order);
        PYR_CHECK_LAUNCH();
      });
    }
    int h = 0;
    PYR_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, stream()));
    PYR_HIP(hipStreamSynchronize(stream()));
    if (h)
      throw Error(PYR_ERR_INTERNAL,
                  "adjoint recompute diverged from the forward pass (checkpoints do not match "
                  "the static fields or the particle data)");
  });
}

// FP32C push interpolation on the matrix cores, per mesh: 0 off (default: the
// VALU contraction measured faster inside the register-bound push), 1 on, 2 on
// and count the particle-stages that take the path
int pyr_push_set_mfma(pyr_mesh *mesh, int mode) {
  return guarded([&] {
    PYR_REQUIRE(mesh, "null mesh");
    PYR_REQUIRE(mode >= 0 && mode <= 2, "mode: 0 off, 1 on, 2 on and counting");
    if (mode == 2 && mesh->push_mfma != 2) {
      mesh->mf_count.reserve(sizeof(unsigned long long));
      PYR_HIP(hipMemsetAsync(mesh->mf_count.ptr, 0, sizeof(unsigned long long), stream()));
    }
    mesh->push_mfma = mode;
  });
}

int pyr_push_mfma_waves(pyr_mesh *mesh, int64_t *out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && out, "null argument");
    *out = 0;
    if (!mesh->mf_count.ptr) return;
    unsigned long long v = 0;
    PYR_HIP(hipMemcpyAsync(&v, mesh->mf_count.ptr, sizeof(v), hipMemcpyDeviceToHost, stream()));
    PYR_HIP(hipMemsetAsync(mesh->mf_count.ptr, 0, sizeof(v), stream()));
    PYR_HIP(hipStreamSynchronize(stream()));
    *out = (int64_t)v;
  });
}

int pyr_velocities_backward(int64_t n, int vdim, double c, const double *p, const double *m,
                            const double *v_bar, double *p_bar) {
  return guarded([&] {
    PYR_REQUIRE(n >= 0 && (vdim == 2 || vdim == 3), "invalid argument");
    if (n == 0) return;
    PYR_REQUIRE(p && m && v_bar && p_bar, "null pointer");
    if (vdim == 2)
      hipLaunchKernelGGL(k_velocities_adjoint<2>, dim3(blocks_for(n)), dim3(256), 0, stream(), n,
                         c, p, m, v_bar, p_bar);
    else
      hipLaunchKernelGGL(k_velocities_adjoint<3>, dim3(blocks_for(n)), dim3(256), 0, stream(), n,
                         c, p, m, v_bar, p_bar);
    PYR_CHECK_LAUNCH();
  });
}

}  // extern "C"
