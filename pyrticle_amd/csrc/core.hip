// core.hip -- library state, mesh upload, scans, localization and the
// reference's kill/compaction order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "pyr_internal.h"

namespace pyr {

static thread_local std::string g_last_error;
static hipStream_t g_stream = nullptr;

void set_error(const std::string &msg) { g_last_error = msg; }
hipStream_t stream() { return g_stream; }

// ---------------------------------------------------------------------------
// region timing with HIP events
// ---------------------------------------------------------------------------
static bool g_profiling = false;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof[PROF_NREGIONS];
static hipEvent_t g_prof_open[PROF_NREGIONS];

bool profiling() { return g_profiling; }

void prof_record(int region, bool begin) {
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return;
  (void)hipEventRecord(ev, g_stream);
  if (begin) {
    g_prof_open[region] = ev;
  } else {
    g_prof[region].emplace_back(g_prof_open[region], ev);
  }
}

static void prof_clear() {
  for (auto &v : g_prof) {
    for (auto &pr : v) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    v.clear();
  }
}

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (tiles of 256 threads x 8 items)
// ---------------------------------------------------------------------------
constexpr int SCAN_BS = 256, SCAN_IPT = 8, SCAN_TILE = SCAN_BS * SCAN_IPT;

template <class T>
__global__ __launch_bounds__(SCAN_BS) void k_scan_tiles(const T *in, uint32_t *out, int64_t n,
                                                        uint32_t *tile_sums) {
  __shared__ uint32_t s[SCAN_BS];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)tid * SCAN_IPT;
  uint32_t v[SCAN_IPT];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    int64_t idx = base + k;
    v[k] = idx < n ? (uint32_t)in[idx] : 0u;
    sum += v[k];
  }
  s[tid] = sum;
  __syncthreads();
  for (int off = 1; off < SCAN_BS; off <<= 1) {
    uint32_t t = tid >= off ? s[tid - off] : 0u;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  uint32_t excl = s[tid] - sum;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    int64_t idx = base + k;
    if (idx < n) out[idx] = excl;
    excl += v[k];
  }
  if (tid == SCAN_BS - 1 && tile_sums) tile_sums[blockIdx.x] = s[SCAN_BS - 1];
}

__global__ void k_scan_add(uint32_t *out, int64_t n, const uint32_t *tile_offsets) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += tile_offsets[idx / SCAN_TILE];
}

__global__ void k_copy_u32(const uint32_t *a, uint32_t *b) { *b = *a; }

size_t scan_tmp_words(int64_t n) {
  size_t words = 0;
  while (n > SCAN_TILE) {
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    words += 2 * (size_t)tiles;
    n = tiles;
  }
  return words + 2;
}

template <class T>
static void scan_impl(const T *in, uint32_t *out, int64_t n, uint32_t *total, uint32_t *tmp,
                      hipStream_t s) {
  if (n <= 0) {
    if (total) PYR_HIP(hipMemsetAsync(total, 0, sizeof(uint32_t), s));
    return;
  }
  int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 1) {
    uint32_t *ts = total ? total : tmp;
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3(1), dim3(SCAN_BS), 0, s, in, out, n, ts);
    PYR_CHECK_LAUNCH();
    return;
  }
  uint32_t *sums = tmp, *offs = tmp + tiles;
  hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(SCAN_BS), 0, s, in, out, n,
                     sums);
  PYR_CHECK_LAUNCH();
  scan_impl<uint32_t>(sums, offs, tiles, total, tmp + 2 * tiles, s);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n,
                     offs);
  PYR_CHECK_LAUNCH();
}

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total,
                        uint32_t *tmp, hipStream_t s) {
  scan_impl<uint32_t>(in, out, n, total, tmp, s);
}

static void scan_exclusive_u8(const uint8_t *in, uint32_t *out, int64_t n, uint32_t *total,
                              uint32_t *tmp, hipStream_t s) {
  scan_impl<uint8_t>(in, out, n, total, tmp, s);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <int D>
__global__ void k_find_containing(DevMesh m, int64_t n, const double *__restrict__ x,
                                  uint32_t *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[D], r[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = x[i * D + k];
  out[i] = find_global<D>(m, xi, r);
}

template <int D>
__global__ void k_is_in_element(DevMesh m, int64_t n, const uint32_t *__restrict__ el,
                                const double *__restrict__ x, uint8_t *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi[D], r[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = x[i * D + k];
  uint32_t e = el[i];
  out[i] = (e < (uint32_t)m.K && in_element<D>(m, e, xi, r)) ? 1 : 0;
}

template <int VD>
__global__ void k_velocities(int64_t n, double c, const double *__restrict__ p,
                             const double *__restrict__ mass, double *__restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double pi[VD], vi[VD];
#pragma unroll
  for (int k = 0; k < VD; ++k) pi[k] = p[i * VD + k];
  velocity<VD>(pi, mass[i], c, vi);
#pragma unroll
  for (int k = 0; k < VD; ++k) v[i * VD + k] = vi[k];
}

template <int D>
__global__ void k_locate(DevMesh m, int64_t n, double *__restrict__ x,
                         const double *__restrict__ p, uint32_t *__restrict__ elem,
                         uint8_t *__restrict__ dead, unsigned long long *counters, int *err) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  FindCounters cnt;
  // every lane takes part in the wave's fan and grid searches
  const bool live = i < n;
  double xi[D] = {}, pi[D] = {}, r[D];
  uint32_t e = INVALID;
  if (live) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xi[k] = x[i * D + k];
      pi[k] = p[i * D + k];
    }
    e = elem[i];
  }
  int e_flag = 0;
  const bool alive = relocate_wave<D>(m, live, false, xi, pi, e, cnt, e_flag, r);
  if (live) {
    if (e_flag) atomicOr(err, 1);
#pragma unroll
    for (int k = 0; k < D; ++k) x[i * D + k] = xi[k];
    elem[i] = e;
    dead[i] = alive ? 0 : 1;
  }
  flush_counters(cnt, counters);
}

__global__ void k_holes(int64_t n_new, const uint8_t *__restrict__ dead,
                        const uint32_t *__restrict__ excl, uint32_t *__restrict__ holes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_new && dead[i]) holes[excl[i]] = (uint32_t)i;
}

__global__ void k_donors(int64_t n, int64_t n_new, uint32_t n_dead,
                         const uint8_t *__restrict__ dead, const uint32_t *__restrict__ excl,
                         const uint32_t *__restrict__ holes, uint32_t *__restrict__ from,
                         uint32_t *__restrict__ to) {
  int64_t i = n_new + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || dead[i]) return;
  // alive positions in (i, n) = (n-1-i) - (dead positions in (i, n))
  int64_t rank = (n - 1 - i) - ((int64_t)n_dead - (int64_t)excl[i]);
  from[rank] = (uint32_t)i;
  to[rank] = holes[rank];
}

template <class W>
__global__ void k_apply_moves(W *buf, int64_t words, int64_t n_moves,
                              const uint32_t *__restrict__ from,
                              const uint32_t *__restrict__ to) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t mv = t / words, w = t % words;
  if (mv >= n_moves) return;
  buf[(int64_t)to[mv] * words + w] = buf[(int64_t)from[mv] * words + w];
}

static inline unsigned blocks_for(int64_t n, int bs = 256) {
  return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs);
}

// ---------------------------------------------------------------------------
// mesh upload and the global-search grid
// ---------------------------------------------------------------------------
static void build_grid(pyr_mesh *h, const pyr_mesh_desc *d) {
  const int D = d->dim, K = d->K;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) {
    lo[i] = INFINITY;
    hi[i] = -INFINITY;
  }
  for (int v = 0; v < d->nv; ++v)
    for (int i = 0; i < D; ++i) {
      lo[i] = std::min(lo[i], d->vertices[v * D + i]);
      hi[i] = std::max(hi[i], d->vertices[v * D + i]);
    }
  double ext = 0;
  for (int i = 0; i < D; ++i) ext = std::max(ext, hi[i] - lo[i]);
  const double gmargin = 1e-9 * ext + 1e-300;
  std::vector<double> elo((size_t)K * D), ehi((size_t)K * D);
  double mean_size = 0;
  for (int e = 0; e < K; ++e) {
    double a[3], b[3];
    for (int i = 0; i < D; ++i) {
      a[i] = INFINITY;
      b[i] = -INFINITY;
    }
    for (int k = 0; k <= D; ++k) {
      uint32_t v = d->el_vertices[(size_t)e * (D + 1) + k];
      for (int i = 0; i < D; ++i) {
        a[i] = std::min(a[i], d->vertices[(size_t)v * D + i]);
        b[i] = std::max(b[i], d->vertices[(size_t)v * D + i]);
      }
    }
    double diam = 0;
    for (int i = 0; i < D; ++i) diam = std::max(diam, b[i] - a[i]);
    mean_size += diam;
    const double margin = 1e-6 * diam + gmargin;
    for (int i = 0; i < D; ++i) {
      elo[(size_t)e * D + i] = a[i] - margin;
      ehi[(size_t)e * D + i] = b[i] + margin;
    }
  }
  mean_size /= std::max(1, K);
  DevMesh &m = h->dev;
  long long total = 1;
  for (int i = 0; i < 3; ++i) {
    m.g_lo[i] = 0;
    m.g_inv_h[i] = 1;
    m.g_n[i] = 1;
  }
  double cell = std::max(mean_size * 1.5, 1e-300);
  for (;;) {
    total = 1;
    for (int i = 0; i < D; ++i) {
      double span = (hi[i] - lo[i]) + 4 * gmargin + 2e-6 * ext;
      int nc = std::max(1, (int)std::ceil(span / cell));
      m.g_n[i] = nc;
      total *= nc;
    }
    if (total <= (1LL << 22)) break;
    cell *= 1.5;
  }
  for (int i = 0; i < D; ++i) {
    double span = (hi[i] - lo[i]) + 4 * gmargin + 2e-6 * ext;
    m.g_lo[i] = lo[i] - 2 * gmargin - 1e-6 * ext;
    m.g_inv_h[i] = (double)m.g_n[i] / span;
  }
  std::vector<uint32_t> counts((size_t)total + 1, 0);
  auto cell_range = [&](int e, int i, int &c0, int &c1) {
    double a = (elo[(size_t)e * D + i] - m.g_lo[i]) * m.g_inv_h[i];
    double b = (ehi[(size_t)e * D + i] - m.g_lo[i]) * m.g_inv_h[i];
    c0 = std::max(0, std::min(m.g_n[i] - 1, (int)std::floor(a)));
    c1 = std::max(0, std::min(m.g_n[i] - 1, (int)std::floor(b)));
  };
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<uint32_t> cursor;
    if (pass == 1) {
      for (size_t c = 1; c <= (size_t)total; ++c) counts[c] += counts[c - 1];
      cursor.assign(counts.begin(), counts.end() - 1);
    }
    std::vector<uint32_t> elems(pass == 1 ? counts[total] : 0);
    for (int e = 0; e < K; ++e) {
      int r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0};
      for (int i = 0; i < D; ++i) cell_range(e, i, r0[i], r1[i]);
      for (int c2 = r0[2]; c2 <= r1[2]; ++c2)
        for (int c1 = r0[1]; c1 <= r1[1]; ++c1)
          for (int c0 = r0[0]; c0 <= r1[0]; ++c0) {
            long long cell_id;
            if (D == 3)
              cell_id = ((long long)c2 * m.g_n[1] + c1) * m.g_n[0] + c0;
            else if (D == 2)
              cell_id = (long long)c1 * m.g_n[0] + c0;
            else
              cell_id = c0;
            if (pass == 0)
              counts[cell_id + 1]++;
            else
              elems[cursor[cell_id]++] = (uint32_t)e;
          }
    }
    if (pass == 1) {
      h->owned.emplace_back();
      m.g_starts = upload(h->owned.back(), counts.data(), counts.size());
      h->owned.emplace_back();
      m.g_elems = upload(h->owned.back(), elems.data(), std::max<size_t>(1, elems.size()));
    }
  }
}

}  // namespace pyr

using namespace pyr;

namespace pyr {
void read_step_flags(pyr_mesh *mesh, const int *dev, int out[2]) {
  if (!mesh->h_step) PYR_HIP(hipHostMalloc((void **)&mesh->h_step, 2 * sizeof(int)));
  if (!mesh->ev_step) PYR_HIP(hipEventCreateWithFlags(&mesh->ev_step, hipEventDisableTiming));
  PYR_HIP(hipMemcpyAsync(mesh->h_step, dev, 2 * sizeof(int), hipMemcpyDeviceToHost, stream()));
  PYR_HIP(hipEventRecord(mesh->ev_step, stream()));
  PYR_HIP(hipEventSynchronize(mesh->ev_step));
  out[0] = mesh->h_step[0];
  out[1] = mesh->h_step[1];
}
}  // namespace pyr

pyr_mesh::~pyr_mesh() {
  if (h_step) (void)hipHostFree(h_step);
  if (ev_step) (void)hipEventDestroy(ev_step);
  for (auto &b : owned) b.release();
  minv.release();
  minv32.release();
  exps.release();
  s_modal.release();
  s_flags.release();
  s_scan.release();
  s_tmp.release();
  s_holes.release();
  s_misc.release();
  s_adj.release();
  counters.release();
  sf64.release();
  sf32.release();
  sf_nodal.release();
  s_avg.release();
  s_occ.release();
  mf_count.release();
}

void pyr_deposit_cache_free(pyr_mesh *mesh);   // deposit.hip

// the current device's copy of pyr_compact_plan_static's scratch (below)
static uint32_t *static_scratch();

extern "C" {

int pyr_version(void) { return 1; }

const char *pyr_last_error(void) { return g_last_error.c_str(); }

int pyr_init(int device) {
  return guarded([&] {
    int count = 0;
    PYR_HIP(hipGetDeviceCount(&count));
    PYR_REQUIRE(count > 0, "no HIP device visible");
    PYR_REQUIRE(device >= 0 && device < count, "device index out of range");
    PYR_HIP(hipSetDevice(device));
    static_scratch();   // this device's copy of the static plan scratch
  });
}

int pyr_set_stream(void *s) {
  g_stream = (hipStream_t)s;
  return PYR_OK;
}

int pyr_sync(void) {
  return guarded([&] { PYR_HIP(hipStreamSynchronize(g_stream)); });
}

int pyr_profile_enable(int on) {
  return guarded([&] {
    prof_clear();
    g_profiling = on != 0;
  });
}

int pyr_profile_read(int region, double *total_ms, int64_t *launches) {
  return guarded([&] {
    PYR_REQUIRE(region >= 0 && region < PROF_NREGIONS && total_ms && launches,
                "invalid argument");
    PYR_HIP(hipDeviceSynchronize());
    double t = 0;
    for (auto &pr : g_prof[region]) {
      float ms = 0;
      PYR_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      t += ms;
    }
    *total_ms = t;
    *launches = (int64_t)g_prof[region].size();
  });
}

int pyr_mesh_create(const pyr_mesh_desc *d, pyr_mesh **out) {
  return guarded([&] {
    PYR_REQUIRE(d && out, "null argument");
    PYR_REQUIRE(d->dim == 2 || d->dim == 3, "dim must be 2 or 3");
    PYR_REQUIRE(d->K > 0 && d->Np > 0 && d->nv > 0, "empty mesh");
    PYR_REQUIRE(d->Np <= 64, "at most 64 nodes per element");
    auto *h = new pyr_mesh();
    try {
      const int D = d->dim, K = d->K, F = D + 1;
      h->dim = D;
      h->K = K;
      h->Np = d->Np;
      h->nv = d->nv;
      DevMesh &m = h->dev;
      m.dim = D;
      m.K = K;
      m.Np = d->Np;
      m.nv = d->nv;
      h->owned.reserve(32);
      auto up = [&](const auto *p, size_t count) {
        h->owned.emplace_back();
        return upload(h->owned.back(), p, count);
      };
      m.inv_map = up(d->inv_map, (size_t)K * (D * D + D));
      m.jacobian = d->jacobian ? up(d->jacobian, (size_t)K) : nullptr;
      m.el_vertices = up(d->el_vertices, (size_t)K * F);
      m.face_normal = up(d->face_normal, (size_t)K * F * D);
      m.face_rhs = up(d->face_rhs, (size_t)K * F);
      m.face_centroid = up(d->face_centroid, (size_t)K * F * D);
      m.face_radius = up(d->face_radius, (size_t)K * F);
      m.face_neighbor = up(d->face_neighbor, (size_t)K * F);
      m.face_per_axis = up(d->face_per_axis, (size_t)K * F);
      m.vertices = up(d->vertices, (size_t)d->nv * D);
      m.vadj_starts = up(d->vadj_starts, (size_t)d->nv + 1);
      m.vadj_elements = up(d->vadj_elements, (size_t)std::max(1, d->nadj));
      m.nodes = up(d->nodes, (size_t)K * d->Np * D);
      for (int i = 0; i < 3; ++i) {
        m.per_min[i] = i < D ? d->periodicity[2 * i] : 0;
        m.per_max[i] = i < D ? d->periodicity[2 * i + 1] : 0;
      }
      build_grid(h, d);
      // interpolation: Minv = V^{-1} so that modal coefficients c = Minv f
      // give f(x) = sum_k c_k phi_k(r(x)), the same interpolant as the
      // reference's Lagrange weights l = V^{-T} phi (push_monomial.hpp:192).
      const int np = d->Np;
      std::vector<double> a(d->vandermonde, d->vandermonde + (size_t)np * np);
      std::vector<double> inv((size_t)np * np, 0.0);
      for (int i = 0; i < np; ++i) inv[(size_t)i * np + i] = 1.0;
      for (int c = 0; c < np; ++c) {   // Gauss-Jordan with partial pivoting
        int piv = c;
        for (int r = c + 1; r < np; ++r)
          if (std::fabs(a[(size_t)r * np + c]) > std::fabs(a[(size_t)piv * np + c])) piv = r;
        PYR_REQUIRE(a[(size_t)piv * np + c] != 0.0, "singular Vandermonde matrix");
        if (piv != c)
          for (int k = 0; k < np; ++k) {
            std::swap(a[(size_t)piv * np + k], a[(size_t)c * np + k]);
            std::swap(inv[(size_t)piv * np + k], inv[(size_t)c * np + k]);
          }
        const double s = 1.0 / a[(size_t)c * np + c];
        for (int k = 0; k < np; ++k) {
          a[(size_t)c * np + k] *= s;
          inv[(size_t)c * np + k] *= s;
        }
        for (int r = 0; r < np; ++r)
          if (r != c) {
            const double f = a[(size_t)r * np + c];
            if (f != 0.0)
              for (int k = 0; k < np; ++k) {
                a[(size_t)r * np + k] -= f * a[(size_t)c * np + k];
                inv[(size_t)r * np + k] -= f * inv[(size_t)c * np + k];
              }
          }
      }
      upload(h->minv, inv.data(), inv.size());
      std::vector<float> inv32(inv.begin(), inv.end());
      upload(h->minv32, inv32.data(), inv32.size());
      h->h_exps.assign(d->exponents, d->exponents + (size_t)np * D);
      upload(h->exps, h->h_exps.data(), h->h_exps.size());
      h->counters.reserve(4 * sizeof(unsigned long long));
      PYR_HIP(hipDeviceSynchronize());
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int pyr_mesh_destroy(pyr_mesh *mesh) {
  return guarded([&] {
    if (!mesh) return;
    PYR_HIP(hipDeviceSynchronize());
    pyr_deposit_cache_free(mesh);
    delete mesh;
  });
}

int pyr_find_containing_element(pyr_mesh *mesh, int64_t n, const double *x,
                                uint32_t *elem_out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n >= 0, "invalid argument");
    if (n == 0) return;
    PYR_REQUIRE(x && elem_out, "null pointer");
    if (mesh->dim == 2)
      hipLaunchKernelGGL(k_find_containing<2>, dim3(blocks_for(n)), dim3(256), 0, stream(),
                         mesh->dev, n, x, elem_out);
    else
      hipLaunchKernelGGL(k_find_containing<3>, dim3(blocks_for(n)), dim3(256), 0, stream(),
                         mesh->dev, n, x, elem_out);
    PYR_CHECK_LAUNCH();
  });
}

int pyr_is_in_element(pyr_mesh *mesh, int64_t n, const uint32_t *elem, const double *x,
                      uint8_t *inside_out) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n >= 0, "invalid argument");
    if (n == 0) return;
    PYR_REQUIRE(elem && x && inside_out, "null pointer");
    if (mesh->dim == 2)
      hipLaunchKernelGGL(k_is_in_element<2>, dim3(blocks_for(n)), dim3(256), 0, stream(),
                         mesh->dev, n, elem, x, inside_out);
    else
      hipLaunchKernelGGL(k_is_in_element<3>, dim3(blocks_for(n)), dim3(256), 0, stream(),
                         mesh->dev, n, elem, x, inside_out);
    PYR_CHECK_LAUNCH();
  });
}

int pyr_velocities(int64_t n, int vdim, double c, const double *p, const double *m,
                   double *v) {
  return guarded([&] {
    PYR_REQUIRE(n >= 0 && (vdim == 2 || vdim == 3), "invalid argument");
    if (n == 0) return;
    PYR_REQUIRE(p && m && v, "null pointer");
    if (vdim == 2)
      hipLaunchKernelGGL(k_velocities<2>, dim3(blocks_for(n)), dim3(256), 0, stream(), n, c,
                         p, m, v);
    else
      hipLaunchKernelGGL(k_velocities<3>, dim3(blocks_for(n)), dim3(256), 0, stream(), n, c,
                         p, m, v);
    PYR_CHECK_LAUNCH();
  });
}

int pyr_locate(pyr_mesh *mesh, int64_t n, double *x, const double *p, uint32_t *elem,
               uint8_t *dead, unsigned long long *counters) {
  return guarded([&] {
    PYR_REQUIRE(mesh && n >= 0, "invalid argument");
    if (n == 0) return;
    PYR_REQUIRE(x && p && elem && dead, "null pointer");
    mesh->s_misc.reserve(64);
    int *err = mesh->s_misc.as<int>();
    PYR_HIP(hipMemsetAsync(err, 0, sizeof(int), stream()));
    if (mesh->dim == 2)
      hipLaunchKernelGGL(k_locate<2>, dim3(blocks_for(n)), dim3(256), 0, stream(), mesh->dev,
                         n, x, p, elem, dead, counters, err);
    else
      hipLaunchKernelGGL(k_locate<3>, dim3(blocks_for(n)), dim3(256), 0, stream(), mesh->dev,
                         n, x, p, elem, dead, counters, err);
    PYR_CHECK_LAUNCH();
    int h_err = 0;
    PYR_HIP(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, stream()));
    PYR_HIP(hipStreamSynchronize(stream()));
    if (h_err)
      throw Error(PYR_ERR_NO_BEST_NORMAL, "no best normal found--weird");
  });
}

// Scratch of pyr_compact_plan (scan of the dead flags, holes, dead count): one
// library-level set, grown on demand, serialized by its mutex, freed by
// pyr_release_scratch.
static struct {
  std::mutex mu;
  DevBuf excl, tmp, holes, small;
} g_plan;

int pyr_release_scratch(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lock(g_plan.mu);
    for (DevBuf *b : {&g_plan.excl, &g_plan.tmp, &g_plan.holes, &g_plan.small}) b->release();
  });
}

}  // extern "C"

// The swap-remove plan of particle_state.hpp:283-293/361-377 from dead flags,
// on caller-provided scratch: excl (n words, the exclusive scan of the flags),
// holes (n words), tmp (scan_tmp_words(n) words) and total (1 word).  Every
// scratch word that is read was written earlier in the same call (the scan
// writes excl and total; k_holes writes holes[0 .. holes below n_new), the
// entries k_donors reads), so nothing needs clearing between calls.
static void compact_plan_on(int64_t n, const uint8_t *dead, uint32_t *move_from,
                            uint32_t *move_to, int64_t *n_new_out, int64_t *n_moves_out,
                            uint32_t *excl, uint32_t *holes, uint32_t *tmp, uint32_t *total) {
  scan_exclusive_u8(dead, excl, n, total, tmp, stream());
  uint32_t n_dead = 0;
  PYR_HIP(hipMemcpyAsync(&n_dead, total, sizeof(uint32_t), hipMemcpyDeviceToHost, stream()));
  PYR_HIP(hipStreamSynchronize(stream()));
  const int64_t n_new = n - (int64_t)n_dead;
  int64_t n_moves = 0;
  if (n_dead && n_new > 0) {
    uint32_t holes_count = 0;   // dead positions below n_new
    PYR_HIP(hipMemcpyAsync(&holes_count, excl + n_new, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           stream()));
    hipLaunchKernelGGL(k_holes, dim3(blocks_for(n_new)), dim3(256), 0, stream(), n_new, dead,
                       excl, holes);
    PYR_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_donors, dim3(blocks_for(n - n_new)), dim3(256), 0, stream(), n, n_new,
                       n_dead, dead, excl, holes, move_from, move_to);
    PYR_CHECK_LAUNCH();
    PYR_HIP(hipStreamSynchronize(stream()));
    n_moves = holes_count;
  }
  *n_new_out = n_new;
  *n_moves_out = n_moves;
}

// Scratch of pyr_compact_plan_static, in the library image of every device:
// 64 words (the dead count at word 0), then excl and holes (PYR_STATIC_PLAN_MAX
// words each), then the scan's tile partials.  Host code uses the per-device
// address from hipGetSymbolAddress, never the symbol itself.
constexpr int64_t STATIC_TMP_WORDS = 2048;
constexpr int64_t STATIC_WORDS = 64 + 2 * (int64_t)PYR_STATIC_PLAN_MAX + STATIC_TMP_WORDS;
static __device__ __align__(256) uint32_t scratch[STATIC_WORDS];
static_assert(2 * ((PYR_STATIC_PLAN_MAX + 2047) / 2048) + 2 <= STATIC_TMP_WORDS,
              "the scan partials of the largest plan must fit");

static std::mutex g_static_mu;
static uint32_t *g_static_ptr[64] = {};   // per device

// Resolves the current device's copy of the scratch array (once per device).
static uint32_t *static_scratch() {
  int dev = 0;
  PYR_HIP(hipGetDevice(&dev));
  PYR_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void *p = nullptr;
    PYR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PYR_REQUIRE(p, "static scratch has no device address");
    g_static_ptr[dev] = static_cast<uint32_t *>(p);
  }
  return g_static_ptr[dev];
}

extern "C" {

int pyr_compact_plan(int64_t n, const uint8_t *dead, uint32_t *move_from, uint32_t *move_to,
                     int64_t *n_new_out, int64_t *n_moves_out) {
  return guarded([&] {
    PYR_REQUIRE(n >= 0 && n < (int64_t)0xFFFFFFFF, "invalid particle count");
    PYR_REQUIRE(n_new_out && n_moves_out, "null pointer");
    if (n == 0) {
      *n_new_out = 0;
      *n_moves_out = 0;
      return;
    }
    PYR_REQUIRE(dead && move_from && move_to, "null pointer");
    std::lock_guard<std::mutex> lock(g_plan.mu);
    g_plan.excl.reserve(sizeof(uint32_t) * (size_t)n);
    g_plan.tmp.reserve(sizeof(uint32_t) * scan_tmp_words(n));
    g_plan.holes.reserve(sizeof(uint32_t) * (size_t)n);
    g_plan.small.reserve(2 * sizeof(uint32_t));
    compact_plan_on(n, dead, move_from, move_to, n_new_out, n_moves_out,
                    g_plan.excl.as<uint32_t>(), g_plan.holes.as<uint32_t>(),
                    g_plan.tmp.as<uint32_t>(), g_plan.small.as<uint32_t>());
  });
}

int pyr_compact_plan_static(int64_t n, const uint8_t *dead, uint32_t *move_from,
                            uint32_t *move_to, int64_t *n_new_out, int64_t *n_moves_out) {
  return guarded([&] {
    PYR_REQUIRE(n >= 0, "invalid particle count");
    PYR_REQUIRE(n <= (int64_t)PYR_STATIC_PLAN_MAX,
                "pyr_compact_plan_static handles at most PYR_STATIC_PLAN_MAX (1048576) "
                "particles; use pyr_compact_plan for larger counts");
    PYR_REQUIRE(n_new_out && n_moves_out, "null pointer");
    uint32_t *base = static_scratch();
    if (n == 0) {
      *n_new_out = 0;
      *n_moves_out = 0;
      return;
    }
    PYR_REQUIRE(dead && move_from && move_to, "null pointer");
    PYR_REQUIRE((int64_t)scan_tmp_words(n) <= STATIC_TMP_WORDS, "scan partials exceed scratch");
    compact_plan_on(n, dead, move_from, move_to, n_new_out, n_moves_out, base + 64,
                    base + 64 + PYR_STATIC_PLAN_MAX, base + 64 + 2 * PYR_STATIC_PLAN_MAX,
                    base);
  });
}

int pyr_apply_moves(void *buf, int64_t elem_bytes, int64_t n_moves, const uint32_t *from,
                    const uint32_t *to) {
  return guarded([&] {
    PYR_REQUIRE(elem_bytes > 0 && n_moves >= 0, "invalid argument");
    if (n_moves == 0) return;
    PYR_REQUIRE(buf && from && to, "null pointer");
    if (elem_bytes % 8 == 0) {
      int64_t w = elem_bytes / 8;
      hipLaunchKernelGGL(k_apply_moves<uint64_t>, dim3(blocks_for(n_moves * w)), dim3(256),
                         0, stream(), (uint64_t *)buf, w, n_moves, from, to);
    } else if (elem_bytes % 4 == 0) {
      int64_t w = elem_bytes / 4;
      hipLaunchKernelGGL(k_apply_moves<uint32_t>, dim3(blocks_for(n_moves * w)), dim3(256),
                         0, stream(), (uint32_t *)buf, w, n_moves, from, to);
    } else {
      hipLaunchKernelGGL(k_apply_moves<uint8_t>, dim3(blocks_for(n_moves * elem_bytes)),
                         dim3(256), 0, stream(), (uint8_t *)buf, elem_bytes, n_moves, from,
                         to);
    }
    PYR_CHECK_LAUNCH();
  });
}

}  // extern "C"
