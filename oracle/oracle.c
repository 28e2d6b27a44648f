/*
 * oracle.c -- CPU restatement of pyrticle's particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, the smoke check and the benchmark's CPU-baseline
 * leg use it, as the checker.  It is single-threaded FP64 C, compiled with
 * -ffp-contract=off so that every a*b+c rounds twice, as the reference C++
 * (g++, no FMA contraction on x86-64 by default) does.
 *
 * Each function names the reference lines it restates.  Parts that the
 * reference delegates to third-party code are restated from the published
 * algorithm:
 *   - hedge affine_map::operator() = prod(A, x) + b (ublas row-wise dot);
 *   - LAPACK dgetf2/dgetrs and reference-BLAS dtrsm/dger/dlaswp (netlib);
 *   - Boost.Math beta(a, b) = Gamma(a) Gamma(b) / Gamma(a + b);
 *   - hedge LSRK4TimeStepper: Carpenter & Kennedy (1994), NASA TM-109112,
 *     five-stage fourth-order 2N-storage coefficients.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INVALID 0xFFFFFFFFu

typedef struct {
  int dim, K, Np, nv;
  const double *inv_map;        /* K x (d*d + d): A row-major, then b */
  const double *jacobian;       /* K */
  const uint32_t *el_vertices;  /* K x (d+1) */
  const double *face_normal;    /* K x (d+1) x d */
  const double *face_rhs;       /* K x (d+1) */
  const double *face_centroid;  /* K x (d+1) x d */
  const double *face_radius;    /* K x (d+1) */
  const uint32_t *face_neighbor;/* K x (d+1) */
  const uint32_t *face_per_axis;/* K x (d+1) */
  const double *vertices;       /* nv x d */
  const uint32_t *vadj_starts;  /* nv + 1 */
  const uint32_t *vadj_elements;
  const double *periodicity;    /* d x 2: (min, max); min == max: not periodic */
  const double *nodes;          /* K*Np x d */
} orc_mesh;

typedef struct {
  uint64_t same, neighbor, vertex, global;
} orc_counters;

static char g_err[512];
const char *orc_last_error(void) { return g_err; }

/* ---------------------------------------------------------------------- */
/* meshdata.hpp:38-48 is_in_unit_simplex                                   */
int orc_is_in_unit_simplex(int d, const double *r, double tol) {
  for (int i = 0; i < d; ++i)
    if (r[i] < -1 - tol) return 0;
  double s = 0;
  for (int i = 0; i < d; ++i) s += r[i];
  return s <= (double)(-(d - 2)) + tol;
}

/* hedge affine_map<double>::operator(): prod(matrix, x) + vector           */
void orc_inverse_map(const orc_mesh *m, uint32_t en, const double *x, double *r) {
  const int d = m->dim;
  const double *a = m->inv_map + (size_t)en * (d * d + d);
  const double *b = a + d * d;
  for (int i = 0; i < d; ++i) {
    double t = 0;
    for (int j = 0; j < d; ++j) t += a[i * d + j] * x[j];
    r[i] = t + b[i];
  }
}

static int in_element(const orc_mesh *m, uint32_t en, const double *pt) {
  double r[3];
  orc_inverse_map(m, en, pt, r);
  return orc_is_in_unit_simplex(m->dim, r, 1e-10);
}

/* meshdata.hpp:208-215 find_containing_element (first in index order)      */
uint32_t orc_find_containing_element(const orc_mesh *m, const double *pt) {
  for (int en = 0; en < m->K; ++en)
    if (in_element(m, (uint32_t)en, pt)) return (uint32_t)en;
  return ORC_INVALID;
}

/* particle_state.hpp:130-261 find_new_containing_element.
 * Returns the element, ORC_INVALID, or sets *err when every face normal has
 * n.p <= 0 (the reference throws "no best normal found--weird").           */
uint32_t orc_find_new_containing_element(const orc_mesh *m, const double *pt,
                                         const double *mom, uint32_t prev,
                                         orc_counters *c, int *err) {
  const int d = m->dim;
  if (prev != ORC_INVALID) {
    if (in_element(m, prev, pt)) {
      c->same++;
      return prev;
    }
    {
      int closest = -1;
      double max_ip = 0;
      for (int f = 0; f <= d; ++f) {
        const double *n = m->face_normal + ((size_t)prev * (d + 1) + f) * d;
        double ip = 0;
        for (int i = 0; i < d; ++i) ip += n[i] * mom[i];
        if (ip > max_ip) {
          closest = f;
          max_ip = ip;
        }
      }
      if (closest == -1) {
        snprintf(g_err, sizeof g_err, "no best normal found--weird (momentum=[%g,%g,%g])",
                 mom[0], d > 1 ? mom[1] : 0.0, d > 2 ? mom[2] : 0.0);
        *err = 1;
        return ORC_INVALID;
      }
      uint32_t possible = m->face_neighbor[(size_t)prev * (d + 1) + closest];
      if (possible != ORC_INVALID && in_element(m, possible, pt)) {
        c->neighbor++;
        return possible;
      }
    }
    {
      uint32_t closest_vertex = ORC_INVALID;
      double min_dist = INFINITY;
      for (int k = 0; k <= d; ++k) {
        uint32_t vi = m->el_vertices[(size_t)prev * (d + 1) + k];
        double s = 0;
        for (int i = 0; i < d; ++i) {
          double dv = m->vertices[(size_t)vi * d + i] - pt[i];
          s += dv * dv;
        }
        double dist = sqrt(s);
        if (dist < min_dist) {
          closest_vertex = vi;
          min_dist = dist;
        }
      }
      for (uint32_t k = m->vadj_starts[closest_vertex];
           k < m->vadj_starts[closest_vertex + 1]; ++k) {
        uint32_t possible = m->vadj_elements[k];
        if (in_element(m, possible, pt)) {
          c->vertex++;
          return possible;
        }
      }
    }
  }
  c->global++;
  return orc_find_containing_element(m, pt);
}

/* ---------------------------------------------------------------------- */
/* particle state with extra per-particle vectors that follow moves
 * (pyrticle/tools.py:106-178 NumberShiftableVector)                         */
typedef struct {
  int xd, vd;
  int64_t n;
  double *x, *p, *q, *m;
  uint32_t *elem;
  int nextra;
  double **extra;    /* each extra[k] has width extra_w[k] doubles/particle */
  const int *extra_w;
  int64_t *ids;      /* optional identity array, moved along */
} orc_state;

/* particle_state.hpp:386-407 move_particle                                 */
static void move_particle(orc_state *s, int64_t from, int64_t to) {
  s->elem[to] = s->elem[from];
  for (int k = 0; k < s->nextra; ++k) {
    int w = s->extra_w[k];
    for (int i = 0; i < w; ++i) s->extra[k][to * w + i] = s->extra[k][from * w + i];
  }
  if (s->ids) s->ids[to] = s->ids[from];
  for (int i = 0; i < s->xd; ++i) s->x[to * s->xd + i] = s->x[from * s->xd + i];
  for (int i = 0; i < s->vd; ++i) s->p[to * s->vd + i] = s->p[from * s->vd + i];
  s->q[to] = s->q[from];
  s->m[to] = s->m[from];
}

/* particle_state.hpp:361-377 kill_particle                                 */
static void kill_particle(orc_state *s, int64_t pn) {
  --s->n;
  if (s->n) move_particle(s, s->n, pn);
}

/* particle_state.hpp:301-357 boundary_hit; returns 1 if the particle was
 * re-found after a periodic wrap, 0 if it must die.                        */
static int boundary_hit(const orc_mesh *m, orc_state *s, int64_t pn, orc_counters *c,
                        int *err) {
  const int d = m->dim;
  double pt[3];
  int trip = 0;
  for (int i = 0; i < d; ++i) pt[i] = s->x[pn * s->xd + i];
  for (int axis = 0; axis < d; ++axis) {
    double mn = m->periodicity[2 * axis], mx = m->periodicity[2 * axis + 1];
    if (mn != mx) {
      double xi = pt[axis];
      if (xi < mn) {
        pt[axis] += (mx - mn);
        trip = 1;
      } else if (xi > mx) {
        pt[axis] -= (mx - mn);
        trip = 1;
      }
    }
  }
  if (trip) {
    for (int i = 0; i < d; ++i) s->x[pn * s->xd + i] = pt[i];
    uint32_t ce = orc_find_new_containing_element(m, s->x + pn * s->xd, s->p + pn * s->vd,
                                                  s->elem[pn], c, err);
    if (*err) return 0;
    if (ce != ORC_INVALID) {
      s->elem[pn] = ce;
      return 1;
    }
  }
  return 0;
}

/* particle_state.hpp:266-295 update_containing_elements, with the kill
 * listener of cloud.py:581-585 (kill_particle on every boundary hit).      */
static int update_containing_elements(const orc_mesh *m, orc_state *s, orc_counters *c) {
  int err = 0;
  for (int64_t pn = 0; pn < s->n;) {
    uint32_t prev = s->elem[pn];
    uint32_t ne = orc_find_new_containing_element(m, s->x + pn * s->xd, s->p + pn * s->vd,
                                                  prev, c, &err);
    if (err) return -1;
    if (ne == ORC_INVALID) {
      int found = boundary_hit(m, s, pn, c, &err);
      if (err) return -1;
      if (!found) kill_particle(s, pn);
      /* a re-found particle keeps pn and is examined again, as the
         reference's loop does (pn does not advance) */
    } else {
      s->elem[pn] = ne;
      ++pn;
    }
  }
  return 0;
}

int orc_update_containing_elements(const orc_mesh *m, int xd, int vd, int64_t *n,
                                   double *x, double *p, double *q, double *mass,
                                   uint32_t *elem, int nextra, double **extra,
                                   const int *extra_w, int64_t *ids, orc_counters *c) {
  orc_state s = {xd, vd, *n, x, p, q, mass, elem, nextra, extra, extra_w, ids};
  int rc = update_containing_elements(m, &s, c);
  *n = s.n;
  return rc;
}

/* ---------------------------------------------------------------------- */
/* particle_state.hpp:98-125 get_velocities                                 */
void orc_velocities(int64_t n, int vd, double c, const double *p, const double *mass,
                    double *v) {
  for (int64_t pn = 0; pn < n; ++pn) {
    const double m = mass[pn];
    double s = 0;
    for (int i = 0; i < vd; ++i) s += p[pn * vd + i] * p[pn * vd + i];
    double pm = sqrt(s);
    if (pm == 0) {
      for (int i = 0; i < vd; ++i) v[pn * vd + i] = 0;
    } else {
      double vm = c * pm / sqrt(m * m * c * c + pm * pm);
      double f = vm / pm;
      for (int i = 0; i < vd; ++i) v[pn * vd + i] = f * p[pn * vd + i];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* tools.cpp:35-58 sphere_area and the polynomial shape normalizer           */
double orc_shape_normalizer(double radius, int dims, double alpha) {
  const double n = dims;
  double sphere_area = 2 * pow(M_PI, n / 2) / tgamma(n / 2);
  double a = n / 2, b = alpha + 1;
  double beta = tgamma(a) * tgamma(b) / tgamma(a + b);
  return 2 / (sphere_area * pow(radius, n + alpha) * beta);
}

/* tools.hpp:520-534 polynomial_shape_function::operator()                   */
static inline double shape_eval(double normalizer, double radius, double alpha,
                                const double *r, int d) {
  double r2 = 0;
  for (int i = 0; i < d; ++i) r2 += r[i] * r[i];
  if (r2 > radius * radius) return 0;
  double t = radius - r2 / radius;
  if (alpha == 2) return normalizer * t * t;
  return normalizer * pow(t, alpha);
}

double orc_shape_value(double normalizer, double radius, double alpha, int d,
                       const double *r) {
  return shape_eval(normalizer, radius, alpha, r, d);
}

/* ---------------------------------------------------------------------- */
/* element_finder.hpp:203-294 face_based_element_finder, feeding
 * dep_shape.hpp:88-106 element_target and dep_target.hpp:54-186 targets.   */
typedef struct {
  const orc_mesh *m;
  double normalizer, radius, alpha;
  double charge;
  const double *vel;   /* vd entries or NULL */
  int vd;
  double *rho, *j;
  int64_t *stamp;
  int64_t mark;
  uint64_t visits, elements;
  int depth, max_depth;
  /* normalized deposition (dep_normshape.hpp:88-174): record instead of add */
  const double *iw;       /* integral weights M 1 (Np), or NULL */
  double integral;        /* sum over visited elements of |J| sum_i S_i w_i */
  double *rec_s;          /* recorded shape values */
  uint32_t *rec_el;       /* recorded elements */
  size_t rec_n, rec_cap;  /* recorded elements, capacity in elements */
  /* walk statistics (dep_normshape.hpp:131-136): visits and centroid distances */
  int walk;
  uint64_t w_n;
  double w_sum, w_sq, w_min, w_max;
} dfs_ctx;

/* dep_normshape.hpp:124-165 add_shape_on_element: store the element's shape
 * values and add |J| sum_i S_i w_i to the particle's integral. */
static int record_shape_on_element(dfs_ctx *c, const double *center, uint32_t en) {
  const orc_mesh *m = c->m;
  const int d = m->dim, np = m->Np;
  const size_t start = (size_t)en * np;
  if (c->rec_n == c->rec_cap) {
    size_t cap = c->rec_cap ? 2 * c->rec_cap : 64;
    double *s2 = (double *)realloc(c->rec_s, sizeof(double) * cap * np);
    uint32_t *e2 = (uint32_t *)realloc(c->rec_el, sizeof(uint32_t) * cap);
    if (!s2 || !e2) return -1;
    c->rec_s = s2;
    c->rec_el = e2;
    c->rec_cap = cap;
  }
  c->elements++;
  double el_integral = 0;
  double *sv = c->rec_s + c->rec_n * np;
  for (int i = 0; i < np; ++i) {
    double r[3];
    for (int k = 0; k < d; ++k) r[k] = m->nodes[(start + i) * d + k] - center[k];
    sv[i] = shape_eval(c->normalizer, c->radius, c->alpha, r, d);
    c->visits++;
    el_integral += sv[i] * c->iw[i];
  }
  c->integral += el_integral * m->jacobian[en];
  c->rec_el[c->rec_n++] = en;
  return 0;
}

/* meshdata.hpp:169-180 element_centroid: the vertex average               */
static void element_centroid(const orc_mesh *m, uint32_t en, double *out) {
  const int d = m->dim;
  for (int i = 0; i < d; ++i) out[i] = m->vertices[(size_t)m->el_vertices[(size_t)en * (d + 1)] * d + i];
  for (int k = 1; k <= d; ++k)
    for (int i = 0; i < d; ++i)
      out[i] += m->vertices[(size_t)m->el_vertices[(size_t)en * (d + 1) + k] * d + i];
  for (int i = 0; i < d; ++i) out[i] /= d + 1;
}

static void add_shape_on_element(dfs_ctx *c, const double *center, uint32_t en) {
  const orc_mesh *m = c->m;
  const int d = m->dim, np = m->Np;
  const size_t start = (size_t)en * np;
  if (c->walk) {
    double cen[3], s = 0;
    element_centroid(m, en, cen);
    for (int i = 0; i < d; ++i) s += (cen[i] - center[i]) * (cen[i] - center[i]);
    const double dist = sqrt(s);
    if (c->w_n == 0 || dist < c->w_min) c->w_min = dist;
    if (c->w_n == 0 || dist > c->w_max) c->w_max = dist;
    c->w_n++;
    c->w_sum += dist;
    c->w_sq += dist * dist;
    return;
  }
  if (c->iw) {
    if (record_shape_on_element(c, center, en)) c->depth = -1000000;   /* out of memory */
    return;
  }
  c->elements++;
  for (int i = 0; i < np; ++i) {
    double r[3];
    for (int k = 0; k < d; ++k) r[k] = m->nodes[(start + i) * d + k] - center[k];
    double el_rho = c->charge * shape_eval(c->normalizer, c->radius, c->alpha, r, d);
    c->visits++;
    if (c->rho) c->rho[start + i] += el_rho;
    if (c->j)
      for (int a = 0; a < c->vd; ++a) c->j[(start + i) * c->vd + a] += c->vel[a] * el_rho;
  }
}

static void recurse(dfs_ctx *c, const double *pos, uint32_t en) {
  const orc_mesh *m = c->m;
  const int d = m->dim;
  if (++c->depth > c->max_depth) c->max_depth = c->depth;
  add_shape_on_element(c, pos, en);
  c->stamp[en] = c->mark;
  for (int f = 0; f <= d; ++f) {
    const size_t fi = (size_t)en * (d + 1) + f;
    uint32_t nb = m->face_neighbor[fi];
    if (nb == ORC_INVALID) continue;
    if (c->stamp[nb] == c->mark) continue;
    const double *n = m->face_normal + fi * d;
    double ip = 0;
    for (int i = 0; i < d; ++i) ip += pos[i] * n[i];
    double plane_dist = fabs(ip - m->face_rhs[fi]);
    if (plane_dist > c->radius) continue;
    const double *fc = m->face_centroid + fi * d;
    double s = 0;
    for (int i = 0; i < d; ++i) {
      double dv = fc[i] - pos[i];
      s += dv * dv;
    }
    if (sqrt(s) > c->radius + m->face_radius[fi]) continue;
    uint32_t axis = m->face_per_axis[fi];
    if (axis == ORC_INVALID) {
      recurse(c, pos, nb);
    } else {
      double pos2[3];
      for (int i = 0; i < d; ++i) pos2[i] = pos[i];
      double mn = m->periodicity[2 * axis], mx = m->periodicity[2 * axis + 1];
      if (pos[axis] - c->radius < mn) {
        pos2[axis] += (mx - mn);
        recurse(c, pos2, nb);
      }
      if (pos[axis] + c->radius > mx) {
        pos2[axis] -= (mx - mn);
        recurse(c, pos2, nb);
      }
    }
  }
  --c->depth;
}

/* dep_target.hpp:203-273 deposit_densities / deposit_j / deposit_rho over
 * the whole particle slice (dep_shape.hpp:128-149).  rho and/or j may be
 * NULL.  visits_out receives the node-visit and element-visit counts.       */
int orc_deposit(const orc_mesh *m, double radius, double alpha, int64_t n, int xd, int vd,
                const double *x, const double *q, const double *v, const uint32_t *elem,
                double *rho, double *j, uint64_t *visits_out) {
  const size_t nn = (size_t)m->K * m->Np;
  if (rho) memset(rho, 0, nn * sizeof(double));
  if (j) memset(j, 0, nn * vd * sizeof(double));
  dfs_ctx c;
  memset(&c, 0, sizeof c);
  c.m = m;
  c.radius = radius;
  c.alpha = alpha;
  c.normalizer = orc_shape_normalizer(radius, m->dim, alpha);
  c.vd = vd;
  c.rho = rho;
  c.j = j;
  c.stamp = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->K);
  if (!c.stamp) return -1;
  for (int k = 0; k < m->K; ++k) c.stamp[k] = -1;
  for (int64_t pn = 0; pn < n; ++pn) {
    double pos[3];
    for (int i = 0; i < m->dim; ++i) pos[i] = x[pn * xd + i];
    c.charge = q[pn];
    c.vel = v ? v + pn * vd : NULL;
    c.mark = pn;
    recurse(&c, pos, elem[pn]);
  }
  free(c.stamp);
  if (visits_out) {
    visits_out[0] = c.visits;
    visits_out[1] = c.elements;
    visits_out[2] = (uint64_t)c.max_depth;
  }
  return 0;
}

/* dep_normshape.hpp:53-245 normalized_shape_function_depositor over the whole
 * particle slice: per particle, the face walk records the shape values of
 * every visited element and the integral I = sum_e |J_e| sum_i S_i w_i
 * (w = M 1, dep_normshape.hpp:214-222); the particle deposits (q / I) S_i,
 * and (q / I) v S_i into j (end_particle, :168-198).  A particle with I = 0
 * deposits nothing and is counted in *n_zero (the reference's warning).
 * scale_out (n, may be NULL) receives q / I (0 where I = 0).                  */
int orc_deposit_normalized(const orc_mesh *m, double radius, double alpha, int64_t n, int xd,
                           int vd, const double *x, const double *q, const double *v,
                           const uint32_t *elem, const double *integral_weights, double *rho,
                           double *j, double *scale_out, int64_t *n_zero) {
  const int np = m->Np;
  const size_t nn = (size_t)m->K * np;
  if (rho) memset(rho, 0, nn * sizeof(double));
  if (j) memset(j, 0, nn * vd * sizeof(double));
  dfs_ctx c;
  memset(&c, 0, sizeof c);
  c.m = m;
  c.radius = radius;
  c.alpha = alpha;
  c.normalizer = orc_shape_normalizer(radius, m->dim, alpha);
  c.vd = vd;
  c.iw = integral_weights;
  c.stamp = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->K);
  if (!c.stamp) return -1;
  for (int k = 0; k < m->K; ++k) c.stamp[k] = -1;
  int64_t zeros = 0;
  int rc = 0;
  for (int64_t pn = 0; pn < n; ++pn) {
    double pos[3];
    for (int i = 0; i < m->dim; ++i) pos[i] = x[pn * xd + i];
    c.mark = pn;
    c.integral = 0;
    c.rec_n = 0;
    c.depth = 0;
    recurse(&c, pos, elem[pn]);
    if (c.depth < 0) {
      rc = -1;
      break;
    }
    if (c.integral == 0) {
      ++zeros;
      if (scale_out) scale_out[pn] = 0;
      continue;
    }
    const double scale = q[pn] / c.integral;
    if (scale_out) scale_out[pn] = scale;
    for (size_t r = 0; r < c.rec_n; ++r) {
      const size_t start = (size_t)c.rec_el[r] * np;
      const double *sv = c.rec_s + r * np;
      for (int i = 0; i < np; ++i) {
        const double val = scale * sv[i];
        if (rho) rho[start + i] += val;
        if (j && v)
          for (int a = 0; a < vd; ++a) j[(start + i) * vd + a] += v[pn * vd + a] * val;
      }
    }
  }
  free(c.stamp);
  free(c.rec_s);
  free(c.rec_el);
  if (n_zero) *n_zero = zeros;
  return rc;
}

/* The face walk of the normalized depositor's statistics (dep_normshape.hpp
 * :131-136 centroid_distance_stats, :175-177 el_per_particle_stats): per
 * particle, the number of elements the walk visits (n_el) and the sum, sum of
 * squares, minimum and maximum of |element centroid - shape centre| over those
 * visits (0 when n_el = 0).                                                   */
int orc_face_walk_stats(const orc_mesh *m, double radius, int64_t n, int xd, const double *x,
                        const uint32_t *elem, uint32_t *n_el, double *dsum, double *dsq,
                        double *dmin, double *dmax) {
  dfs_ctx c;
  memset(&c, 0, sizeof c);
  c.m = m;
  c.radius = radius;
  c.walk = 1;
  c.stamp = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->K);
  if (!c.stamp) return -1;
  for (int k = 0; k < m->K; ++k) c.stamp[k] = -1;
  for (int64_t pn = 0; pn < n; ++pn) {
    double pos[3];
    for (int i = 0; i < m->dim; ++i) pos[i] = x[pn * xd + i];
    c.mark = pn;
    c.w_n = 0;
    c.w_sum = c.w_sq = c.w_min = c.w_max = 0;
    recurse(&c, pos, elem[pn]);
    n_el[pn] = (uint32_t)c.w_n;
    dsum[pn] = c.w_sum;
    dsq[pn] = c.w_sq;
    dmin[pn] = c.w_min;
    dmax[pn] = c.w_max;
  }
  free(c.stamp);
  return 0;
}

/* push_average.hpp:38-443 averaging_particle_pusher::forces with the chained
 * el_/mag_force_averaging_target over the deposition of either depositor
 * (dep_shape.hpp:67-150 when normalized == 0: rho_contrib = q S;
 * dep_normshape.hpp:168-198 when normalized == 1: rho_contrib = (q / I) S and
 * a particle with I = 0 never reaches the targets).  fields[6] (ex..bz,
 * K*Np nodal, NULL = zero_vector).  force (n x vd) = el_force + mag_force.
 * charge_out (n, may be NULL): the reconstructed charge m_particle_charge
 * (0 for particles the targets never saw).  vis_* (all or none):
 * pt_e, pt_b (n x 3), pt_e_stddev, pt_b_stddev (n), el_force, mag_force
 * (n x vd).  n_zero: particles whose targets saw zero charge, plus (normalized)
 * those the depositor skipped.                                                */
int orc_average_forces(const orc_mesh *m, double radius, double alpha, int normalized,
                       int64_t n, int xd, int vd, const double *x, const double *q,
                       const double *v, const uint32_t *elem, const double *integral_weights,
                       const double *const *fields, double *force, double *charge_out,
                       double *vis_e, double *vis_b, double *vis_e_std, double *vis_b_std,
                       double *vis_el, double *vis_mag, int64_t *n_zero) {
  const int np = m->Np;
  dfs_ctx c;
  memset(&c, 0, sizeof c);
  c.m = m;
  c.radius = radius;
  c.alpha = alpha;
  c.normalizer = orc_shape_normalizer(radius, m->dim, alpha);
  c.vd = vd;
  c.iw = integral_weights;
  c.stamp = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->K);
  double *contrib = (double *)malloc(sizeof(double) * np);
  if (!c.stamp || !contrib) return -1;
  for (int k = 0; k < m->K; ++k) c.stamp[k] = -1;
  int64_t zeros = 0;
  int rc = 0;
  for (int64_t pn = 0; pn < n; ++pn) {
    double pos[3];
    for (int i = 0; i < m->dim; ++i) pos[i] = x[pn * xd + i];
    c.mark = pn;
    c.integral = 0;
    c.rec_n = 0;
    c.depth = 0;
    recurse(&c, pos, elem[pn]);
    if (c.depth < 0) {
      rc = -1;
      break;
    }
    for (int a = 0; a < vd; ++a) force[pn * vd + a] = 0;
    if (vis_e) {
      for (int a = 0; a < 3; ++a) vis_e[pn * 3 + a] = vis_b[pn * 3 + a] = 0;
      vis_e_std[pn] = vis_b_std[pn] = 0;
      for (int a = 0; a < vd; ++a) vis_el[pn * vd + a] = vis_mag[pn * vd + a] = 0;
    }
    if (charge_out) charge_out[pn] = 0;
    double dep_scale = q[pn];
    if (normalized) {
      if (c.integral == 0) {   /* dep_normshape.hpp:178-184: targets never see it */
        ++zeros;
        continue;
      }
      dep_scale = q[pn] / c.integral;
    }
    /* begin_particle (:125-131) for both targets; they accumulate alike */
    double qf[2][3] = {{0, 0, 0}, {0, 0, 0}}, sq[2] = {0, 0}, charge = 0;
    double ptf[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (size_t r = 0; r < c.rec_n; ++r) {
      const uint32_t en = c.rec_el[r];
      const size_t start = (size_t)en * np;
      const double *sv = c.rec_s + r * np;
      const double jac = m->jacobian[en];
      for (int i = 0; i < np; ++i) contrib[i] = dep_scale * sv[i] * c.iw[i];
      /* add_shape_on_element (:135-168) */
      for (int h = 0; h < 2; ++h) {
        double qfield[3];
        for (int k = 0; k < 3; ++k) {
          const double *f = fields[3 * h + k];
          double ip = 0;
          if (f)
            for (int i = 0; i < np; ++i) ip += f[start + i] * contrib[i];
          qfield[k] = jac * ip;
          qf[h][k] += qfield[k];
          ptf[h][k] += qfield[k];
        }
        double sqs = 0;
        for (int k = 0; k < 3; ++k) {
          const double *f = fields[3 * h + k];
          double ip = 0;
          if (f)
            for (int i = 0; i < np; ++i) ip += f[start + i] * f[start + i] * contrib[i];
          sqs += ip;
        }
        sq[h] += jac * sqs;
      }
      double ipc = 0;
      for (int i = 0; i < np; ++i) ipc += contrib[i];
      charge += jac * ipc;
    }
    if (charge_out) charge_out[pn] = charge;
    /* end_particle (:170-215, :249-265, :309-330) */
    if (charge == 0) {
      ++zeros;
      continue;
    }
    if (vis_e) {
      for (int h = 0; h < 2; ++h) {
        double *pf = h ? vis_b + pn * 3 : vis_e + pn * 3;
        for (int k = 0; k < 3; ++k) pf[k] = ptf[h][k] / charge;
        const double squared_mean = sq[h] / charge;
        const double mean_squared = pf[0] * pf[0] + pf[1] * pf[1] + pf[2] * pf[2];
        const double variance = squared_mean - mean_squared;
        (h ? vis_b_std : vis_e_std)[pn] = variance < 0 ? 0 : sqrt(variance);
      }
    }
    const double scale = q[pn] / charge;
    double vv[3] = {0, 0, 0};
    for (int a = 0; a < vd; ++a) vv[a] = v[pn * vd + a];
    const double *b = qf[1];
    const double cr[3] = {vv[1] * b[2] - vv[2] * b[1], vv[2] * b[0] - vv[0] * b[2],
                          vv[0] * b[1] - vv[1] * b[0]};
    for (int a = 0; a < vd; ++a) {
      const double el = scale * qf[0][a];
      const double mag = scale * cr[a];
      force[pn * vd + a] = el + mag;
      if (vis_e) {
        vis_el[pn * vd + a] = el;
        vis_mag[pn * vd + a] = mag;
      }
    }
  }
  free(c.stamp);
  free(c.rec_s);
  free(c.rec_el);
  free(contrib);
  if (n_zero) *n_zero = zeros;
  return rc;
}

/* ---------------------------------------------------------------------- */
/* diagnostics.hpp:37-183 (sequential sums in particle order, as there).    */
static double kinetic_energy(const double *p, int vd, double m, double c) {
  double s = 0;
  for (int k = 0; k < vd; ++k) s += p[k] * p[k];
  const double pn = sqrt(s);
  const double v = c * pn / sqrt(m * m * c * c + pn * pn);
  return (pn / v - m) * (c * c);
}

void orc_kinetic_energies(int64_t n, int vd, double c, const double *p, const double *m,
                          double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = kinetic_energy(p + i * vd, vd, m[i], c);
}

double orc_total_charge(int64_t n, const double *q) {
  double r = 0;
  for (int64_t i = 0; i < n; ++i) r += q[i];
  return r;
}

void orc_particle_momentum(int64_t n, int vd, const double *p, double *out) {
  for (int k = 0; k < vd; ++k) out[k] = 0;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < vd; ++k) out[k] += p[i * vd + k];
}

void orc_particle_current(int64_t n, int vd, const double *q, const double *v, double length,
                          double *out) {
  for (int k = 0; k < vd; ++k) out[k] = 0;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < vd; ++k) out[k] += q[i] * v[i * vd + k];
  for (int k = 0; k < vd; ++k) out[k] /= length;
}

double orc_rms_beam_size(int64_t n, int xd, const double *x, int axis) {
  if (n == 0) return 0;
  double r = 0;
  for (int64_t i = 0; i < n; ++i) r += x[i * xd + axis] * x[i * xd + axis];
  return sqrt(r / n);
}

double orc_rms_beam_emittance(int64_t n, int xd, int vd, const double *x, const double *p,
                              int axis, int beam_axis) {
  if (n == 0) return 0;
  double mx = 0, mxp = 0, mx2 = 0, mxp2 = 0, mxxp = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double xi = x[i * xd + axis];
    mx += xi;
    mx2 += xi * xi;
    const double px = p[i * vd + axis], pz = p[i * vd + beam_axis];
    const double xp = pz != 0 ? px / pz : 0;
    mxp += xp;
    mxp2 += xp * xp;
    mxxp += xi * xp;
  }
  mx /= n;
  mxp /= n;
  mx2 /= n;
  mxp2 /= n;
  mxxp /= n;
  return sqrt(mx2 * mxp2 - mx * mx * mxp2 - mx2 * (mxp * mxp) - (mxxp * mxxp - 2 * mxxp * mx * mxp));
}

/* with tools.hpp:303-322 std_dev */
double orc_rms_energy_spread(int64_t n, int vd, double c, const double *p, const double *m) {
  if (n == 0) return 0;
  double sum = 0, sq = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double e = kinetic_energy(p + i * vd, vd, m[i], c);
    sum += e;
    sq += e * e;
  }
  const double mean = sum / n;
  return sqrt(sq / n - mean * mean);
}

/* ---------------------------------------------------------------------- */
/* LAPACK dgetf2 (netlib): LU with partial pivoting of a column-major n x n
 * matrix, as wrap_tools.cpp:126-145 (lu -> getrf) uses it.  ipiv 0-based.  */
int orc_getrf(int n, double *a, int *ipiv) {
  const double sfmin = 2.2250738585072014e-308; /* dlamch('S') */
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int jp = j;
    double best = fabs(a[j + j * n]);
    for (int i = j + 1; i < n; ++i)
      if (fabs(a[i + j * n]) > best) {
        best = fabs(a[i + j * n]);
        jp = i;
      }
    ipiv[j] = jp;
    if (a[jp + j * n] != 0) {
      if (jp != j)
        for (int k = 0; k < n; ++k) {
          double t = a[j + k * n];
          a[j + k * n] = a[jp + k * n];
          a[jp + k * n] = t;
        }
      if (j < n - 1) {
        if (fabs(a[j + j * n]) >= sfmin) {
          double r = 1.0 / a[j + j * n];
          for (int i = j + 1; i < n; ++i) a[i + j * n] *= r;
        } else {
          for (int i = j + 1; i < n; ++i) a[i + j * n] /= a[j + j * n];
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
    if (j < n - 1) {
      /* dger: A(j+1:, j+1:) += x * (-y)^T, x = A(j+1:, j), y = A(j, j+1:) */
      for (int jj = j + 1; jj < n; ++jj) {
        double yv = a[j + jj * n];
        if (yv != 0) {
          double temp = -1.0 * yv;
          for (int i = j + 1; i < n; ++i) a[i + jj * n] += a[i + j * n] * temp;
        }
      }
    }
  }
  return info;
}

/* LAPACK dgetrs('N') (netlib): dlaswp, then dtrsm lower-unit and
 * upper-nonunit, column by column on B (n x nrhs, column-major).           */
void orc_getrs(int n, int nrhs, const double *lu, const int *ipiv, double *b) {
  for (int c = 0; c < nrhs; ++c) {
    double *bc = b + (size_t)c * n;
    for (int i = 0; i < n; ++i)
      if (ipiv[i] != i) {
        double t = bc[i];
        bc[i] = bc[ipiv[i]];
        bc[ipiv[i]] = t;
      }
    for (int k = 0; k < n; ++k)
      if (bc[k] != 0)
        for (int i = k + 1; i < n; ++i) bc[i] -= bc[k] * lu[i + k * n];
    for (int k = n - 1; k >= 0; --k)
      if (bc[k] != 0) {
        bc[k] = bc[k] / lu[k + k * n];
        for (int i = 0; i < k; ++i) bc[i] -= bc[k] * lu[i + k * n];
      }
  }
}

/* ---------------------------------------------------------------------- */
/* push_monomial.hpp:45-70 monomial_basis_function, :155-210 make_interpolator,
 * :99-111 interpolator::operator(), :219-325 forces, tools.hpp:253-264 cross.
 * exps: nb x d exponents; lu/ipiv: LU of V^T (column-major) from orc_getrf.
 * fields[6] = ex, ey, ez, bx, by, bz nodal vectors; NULL = ZeroVector.       */
typedef struct {
  int nb;
  const int *exps;
  const double *lu;
  const int *ipiv;
} orc_interp;

static void lagrange_weights(const orc_mesh *m, const orc_interp *ip, uint32_t en,
                             const double *x, double *coef) {
  double r[3];
  orc_inverse_map(m, en, x, r);
  for (int k = 0; k < ip->nb; ++k) {
    double result = 1;
    for (int i = 0; i < m->dim; ++i) result *= pow(r[i], (double)ip->exps[k * m->dim + i]);
    coef[k] = result;
  }
  orc_getrs(ip->nb, 1, ip->lu, ip->ipiv, coef);
}

int orc_forces(const orc_mesh *m, int nb, const int *exps, const double *lu,
               const int *ipiv, int64_t n, int xd, int vd, const double *x,
               const uint32_t *elem, const double *q, const double *v,
               const double *const *fields, double *force, double *vis_e,
               double *vis_b, double *vis_el, double *vis_mag) {
  orc_interp ip = {nb, exps, lu, ipiv};
  double coef[64];
  if (nb > 64) return -1;
  for (int64_t pn = 0; pn < n; ++pn) {
    uint32_t en = elem[pn];
    lagrange_weights(m, &ip, en, x + pn * xd, coef);
    double f6[6];
    for (int c = 0; c < 6; ++c) {
      if (!fields[c]) {
        f6[c] = 0;
        continue;
      }
      const double *data = fields[c] + (size_t)en * nb;
      double t = 0;
      for (int k = 0; k < nb; ++k) t += coef[k] * data[k];
      f6[c] = t;
    }
    const double charge = q[pn];
    double el[3], qb[3], vv[3], mag[3];
    for (int i = 0; i < 3; ++i) {
      el[i] = charge * f6[i];
      qb[i] = charge * f6[3 + i];
      vv[i] = i < vd ? v[pn * vd + i] : 0;
    }
    mag[0] = vv[1] * qb[2] - vv[2] * qb[1];
    mag[1] = vv[2] * qb[0] - vv[0] * qb[2];
    mag[2] = vv[0] * qb[1] - vv[1] * qb[0];
    for (int i = 0; i < vd; ++i) force[pn * vd + i] = el[i] + mag[i];
    if (vis_e)
      for (int i = 0; i < 3; ++i) {
        vis_e[pn * 3 + i] = f6[i];
        vis_b[pn * 3 + i] = f6[3 + i];
        vis_el[pn * 3 + i] = el[i];
        vis_mag[pn * 3 + i] = mag[i];
      }
  }
  return 0;
}

/* Lagrange weights of one point, for identity checks of the interpolator.  */
void orc_lagrange_weights(const orc_mesh *m, int nb, const int *exps, const double *lu,
                          const int *ipiv, uint32_t en, const double *x, double *coef) {
  orc_interp ip = {nb, exps, lu, ipiv};
  lagrange_weights(m, &ip, en, x, coef);
}

/* ---------------------------------------------------------------------- */
/* hedge LSRK4TimeStepper (Carpenter & Kennedy 1994 coefficients), driving
 * the pyrticle particle RHS (cloud.py:238-311: [v, F, 0]) and advance_state
 * (cloud.py:554-603: x += dx, p += dp, relocalize and kill, with the RK
 * residual vectors following every move, tools.py:106-178).                */
static const double RK4A[5] = {
    0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
    -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double RK4B[5] = {
    1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
    1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
    2277821191437.0 / 14882151754819.0};

void orc_lsrk4_coefficients(double *a, double *b) {
  for (int i = 0; i < 5; ++i) {
    a[i] = RK4A[i];
    b[i] = RK4B[i];
  }
}

int orc_lsrk4_run(const orc_mesh *m, int nb, const int *exps, const double *lu,
                  const int *ipiv, const double *const *fields, double c_light, int xd,
                  int vd, int64_t *n_inout, double *x, double *p, double *q, double *mass,
                  uint32_t *elem, int64_t *ids, double dt, int nsteps, orc_counters *cnt) {
  int64_t n = *n_inout;
  double *rx = (double *)calloc((size_t)(n ? n : 1) * xd, sizeof(double));
  double *rp = (double *)calloc((size_t)(n ? n : 1) * vd, sizeof(double));
  double *v = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1) * vd);
  double *f = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1) * vd);
  if (!rx || !rp || !v || !f) return -1;
  double *extra[2] = {rx, rp};
  int extra_w[2] = {xd, vd};
  int rc = 0;
  for (int step = 0; step < nsteps && rc == 0; ++step) {
    for (int s = 0; s < 5; ++s) {
      orc_velocities(n, vd, c_light, p, mass, v);
      rc = orc_forces(m, nb, exps, lu, ipiv, n, xd, vd, x, elem, q, v, fields, f, NULL,
                      NULL, NULL, NULL);
      if (rc) break;
      const double a = RK4A[s], b = RK4B[s];
      for (int64_t i = 0; i < n * xd; ++i) rx[i] = a * rx[i] + dt * v[i];
      for (int64_t i = 0; i < n * vd; ++i) rp[i] = a * rp[i] + dt * f[i];
      for (int64_t i = 0; i < n * xd; ++i) x[i] += b * rx[i];
      for (int64_t i = 0; i < n * vd; ++i) p[i] += b * rp[i];
      orc_state st = {xd, vd, n, x, p, q, mass, elem, 2, extra, extra_w, ids};
      rc = update_containing_elements(m, &st, cnt);
      n = st.n;
      if (rc) break;
    }
  }
  free(rx);
  free(rp);
  free(v);
  free(f);
  *n_inout = n;
  return rc;
}
