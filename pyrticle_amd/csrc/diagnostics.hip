// diagnostics.hip -- particle-state diagnostics (diagnostics.hpp:37-183,
// bound at wrap_pic.cpp:38-47): kinetic energies, total charge, momentum,
// current, RMS beam size, RMS emittance and RMS energy spread.
//
// The sums run on the device in a fixed order: a fixed grid of RED_BLOCKS
// blocks strides over the particles, each block reduces its threads' partial
// sums in a fixed tree, and one block adds the block partials in a fixed tree.
// The results are therefore the same on every run and differ from the
// reference's sequential sums only in summation order.  The closing formulas
// (square roots, the emittance expansion, the one-pass energy variance) are
// the reference's, evaluated on the host in FP64.
#include "pyr_internal.h"

namespace pyr {

namespace {

constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
constexpr int MAX_Q = 5;

DevBuf g_partials;   // RED_BLOCKS x MAX_Q block partials, then MAX_Q totals

// diagnostics.hpp:37-57 kinetic_energies: E = (p / v - m) c^2 with
// v = c p / sqrt(m^2 c^2 + p^2), p = |momentum| (0 / 0 for a particle at rest,
// as the reference computes it).
__device__ __forceinline__ double kinetic_energy(const double *p, int vdim, double m,
                                                 double c) {
  double s = 0;
  for (int k = 0; k < vdim; ++k) s += p[k] * p[k];
  const double pn = sqrt(s);
  const double v = c * pn / sqrt(m * m * c * c + pn * pn);
  return (pn / v - m) * (c * c);
}

// Per-particle quantities of each reduction.
enum RedKind : int {
  RED_CHARGE = 0,     // q
  RED_MOMENTUM,       // p_k, k < vdim
  RED_CURRENT,        // q v_k, k < vdim
  RED_SIZE,           // x_axis^2
  RED_EMITTANCE,      // x, x', x^2, x'^2, x x'
  RED_ENERGY,         // E, E^2
};

struct RedArgs {
  int kind, xdim, vdim, axis, beam_axis, nq;
  int64_t n;
  double c;
  const double *q, *x, *p, *v, *m;
};

__device__ __forceinline__ void particle_terms(const RedArgs &a, int64_t i, double *t) {
  switch (a.kind) {
    case RED_CHARGE:
      t[0] = a.q[i];
      break;
    case RED_MOMENTUM:
      for (int k = 0; k < a.vdim; ++k) t[k] = a.p[i * a.vdim + k];
      break;
    case RED_CURRENT:
      for (int k = 0; k < a.vdim; ++k) t[k] = a.q[i] * a.v[i * a.vdim + k];
      break;
    case RED_SIZE: {
      const double x = a.x[i * a.xdim + a.axis];
      t[0] = x * x;
      break;
    }
    case RED_EMITTANCE: {
      // diagnostics.hpp:122-136: x' = p_axis / p_beam (0 where p_beam = 0)
      const double x = a.x[i * a.xdim + a.axis];
      const double px = a.p[i * a.vdim + a.axis];
      const double pz = a.p[i * a.vdim + a.beam_axis];
      const double xp = pz != 0.0 ? px / pz : 0.0;
      t[0] = x;
      t[1] = xp;
      t[2] = x * x;
      t[3] = xp * xp;
      t[4] = x * xp;
      break;
    }
    case RED_ENERGY: {
      const double e = kinetic_energy(a.p + i * a.vdim, a.vdim, a.m[i], a.c);
      t[0] = e;
      t[1] = e * e;
      break;
    }
  }
}

// fixed-order tree over the block's threads for each of nq quantities
__device__ void block_tree(double (*sh)[RED_THREADS], int nq) {
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    __syncthreads();
    if ((int)threadIdx.x < w)
      for (int k = 0; k < nq; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
  }
  __syncthreads();
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce(RedArgs a, double *partials) {
  __shared__ double sh[MAX_Q][RED_THREADS];
  double acc[MAX_Q] = {0, 0, 0, 0, 0};
  const int64_t stride = (int64_t)RED_BLOCKS * RED_THREADS;
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < a.n; i += stride) {
    double t[MAX_Q] = {0, 0, 0, 0, 0};
    particle_terms(a, i, t);
#pragma unroll
    for (int k = 0; k < MAX_Q; ++k) acc[k] += t[k];
  }
  for (int k = 0; k < a.nq; ++k) sh[k][threadIdx.x] = acc[k];
  block_tree(sh, a.nq);
  if (threadIdx.x < (unsigned)a.nq) partials[(size_t)blockIdx.x * MAX_Q + threadIdx.x] =
      sh[threadIdx.x][0];
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(int nq, double *partials) {
  __shared__ double sh[MAX_Q][RED_THREADS];
  for (int k = 0; k < nq; ++k) {
    double s = 0;
    for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS)
      s += partials[(size_t)b * MAX_Q + k];
    sh[k][threadIdx.x] = s;
  }
  block_tree(sh, nq);
  if (threadIdx.x < (unsigned)nq)
    partials[(size_t)RED_BLOCKS * MAX_Q + threadIdx.x] = sh[threadIdx.x][0];
}

// runs the reduction and returns the nq sums on the host
void reduce(const RedArgs &a, double *sums) {
  g_partials.reserve(sizeof(double) * ((size_t)RED_BLOCKS * MAX_Q + MAX_Q));
  double *part = g_partials.as<double>();
  hipStream_t s = stream();
  {
    ProfScope prof(PROF_DIAG);
    hipLaunchKernelGGL(k_reduce, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, a, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, s, a.nq, part);
  }
  PYR_CHECK_LAUNCH();
  PYR_HIP(hipMemcpyAsync(sums, part + (size_t)RED_BLOCKS * MAX_Q, sizeof(double) * a.nq,
                         hipMemcpyDeviceToHost, s));
  PYR_HIP(hipStreamSynchronize(s));
}

RedArgs base_args(int kind, int64_t n, int nq) {
  PYR_REQUIRE(n >= 0, "invalid particle count");
  RedArgs a{};
  a.kind = kind;
  a.n = n;
  a.nq = nq;
  return a;
}

__global__ void k_kinetic_energies(int64_t n, int vdim, double c, const double *__restrict__ p,
                                   const double *__restrict__ m, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = kinetic_energy(p + i * vdim, vdim, m[i], c);
}

}  // namespace
}  // namespace pyr

using namespace pyr;

extern "C" {

int pyr_kinetic_energies(int64_t n, int vdim, double vacuum_c, const double *p, const double *m,
                         double *out) {
  return guarded([&] {
    PYR_REQUIRE(n >= 0 && vdim >= 1 && vdim <= 3, "invalid argument");
    if (n == 0) return;
    PYR_REQUIRE(p && m && out, "null pointer");
    hipLaunchKernelGGL(k_kinetic_energies, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream(), n, vdim, vacuum_c, p, m, out);
    PYR_CHECK_LAUNCH();
  });
}

int pyr_total_charge(int64_t n, const double *q, double *out) {
  return guarded([&] {
    PYR_REQUIRE(out, "null pointer");
    RedArgs a = base_args(RED_CHARGE, n, 1);
    *out = 0;
    if (n == 0) return;
    PYR_REQUIRE(q, "null pointer");
    a.q = q;
    reduce(a, out);
  });
}

int pyr_particle_momentum(int64_t n, int vdim, const double *p, double *out) {
  return guarded([&] {
    PYR_REQUIRE(out && vdim >= 1 && vdim <= 3, "invalid argument");
    RedArgs a = base_args(RED_MOMENTUM, n, vdim);
    for (int k = 0; k < vdim; ++k) out[k] = 0;
    if (n == 0) return;
    PYR_REQUIRE(p, "null pointer");
    a.vdim = vdim;
    a.p = p;
    reduce(a, out);
  });
}

int pyr_particle_current(int64_t n, int vdim, const double *q, const double *v, double length,
                         double *out) {
  return guarded([&] {
    PYR_REQUIRE(out && vdim >= 1 && vdim <= 3, "invalid argument");
    RedArgs a = base_args(RED_CURRENT, n, vdim);
    double s[3] = {0, 0, 0};
    if (n > 0) {
      PYR_REQUIRE(q && v, "null pointer");
      a.vdim = vdim;
      a.q = q;
      a.v = v;
      reduce(a, s);
    }
    for (int k = 0; k < vdim; ++k) out[k] = s[k] / length;
  });
}

int pyr_rms_beam_size(int64_t n, int xdim, const double *x, int axis, double *out) {
  return guarded([&] {
    PYR_REQUIRE(out && xdim >= 1 && xdim <= 3, "invalid argument");
    PYR_REQUIRE(axis >= 0 && axis < xdim, "axis out of range");
    RedArgs a = base_args(RED_SIZE, n, 1);
    *out = 0;
    if (n == 0) return;
    PYR_REQUIRE(x, "null pointer");
    a.xdim = xdim;
    a.axis = axis;
    a.x = x;
    double s = 0;
    reduce(a, &s);
    *out = std::sqrt(s / (double)n);
  });
}

int pyr_rms_beam_emittance(int64_t n, int xdim, int vdim, const double *x, const double *p,
                           int axis, int beam_axis, double *out) {
  return guarded([&] {
    PYR_REQUIRE(out && xdim >= 1 && xdim <= 3 && vdim >= 1 && vdim <= 3, "invalid argument");
    PYR_REQUIRE(axis >= 0 && axis < xdim && axis < vdim && beam_axis >= 0 && beam_axis < vdim,
                "axis out of range");
    RedArgs a = base_args(RED_EMITTANCE, n, 5);
    *out = 0;
    if (n == 0) return;
    PYR_REQUIRE(x && p, "null pointer");
    a.xdim = xdim;
    a.vdim = vdim;
    a.axis = axis;
    a.beam_axis = beam_axis;
    a.x = x;
    a.p = p;
    double s[5];
    reduce(a, s);
    // diagnostics.hpp:138-154
    const double cnt = (double)n;
    const double mx = s[0] / cnt, mxp = s[1] / cnt, mx2 = s[2] / cnt, mxp2 = s[3] / cnt,
                 mxxp = s[4] / cnt;
    *out = std::sqrt(mx2 * mxp2 - mx * mx * mxp2 - mx2 * (mxp * mxp) -
                     (mxxp * mxxp - 2 * mxxp * mx * mxp));
  });
}

int pyr_rms_energy_spread(int64_t n, int vdim, double vacuum_c, const double *p,
                          const double *m, double *out) {
  return guarded([&] {
    PYR_REQUIRE(out && vdim >= 1 && vdim <= 3, "invalid argument");
    RedArgs a = base_args(RED_ENERGY, n, 2);
    *out = 0;
    if (n == 0) return;
    PYR_REQUIRE(p && m, "null pointer");
    a.vdim = vdim;
    a.c = vacuum_c;
    a.p = p;
    a.m = m;
    double s[2];
    reduce(a, s);
    // tools.hpp:303-322 std_dev: sqrt(<E^2> - <E>^2)
    const double mean = s[0] / (double)n;
    *out = std::sqrt(s[1] / (double)n - mean * mean);
  });
}

}  // extern "C"
