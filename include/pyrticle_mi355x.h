/*
 * pyrticle_mi355x.h -- C ABI of libpyrticle_mi355x.so, the MI355X (gfx950)
 * implementation of pyrticle's particle hot path: localization, shape-function
 * deposition of rho and j, DG interpolation of E and B, and the relativistic
 * Lorentz push.
 *
 * The reference reaches this path through its Boost.Python module
 * pyrticle._internal (src/wrapper/).  Each entry point below names the
 * reference interface it replaces.  Conventions:
 *
 *   - every function returns 0 (PYR_OK) or a negative PYR_ERR_* code; the
 *     message is available from pyr_last_error() (thread-local).  No C++
 *     exception crosses the ABI.
 *   - particle arrays are DEVICE pointers in the reference's layout:
 *     positions n x xdim and momenta n x vdim (float64, C order), charges and
 *     masses n (float64), containing elements n (uint32).  xdim == vdim == dim.
 *   - work is ordered on the stream given to pyr_set_stream (default: the
 *     null stream); pyr_sync() waits for it.  Functions that return a count
 *     on the host synchronize the stream themselves.
 *   - pyr_mesh_create copies the mesh from HOST memory into HBM; the handle
 *     owns it.
 */
#ifndef PYRTICLE_MI355X_H
#define PYRTICLE_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYR_INVALID_ELEMENT 0xFFFFFFFFu

enum {
  PYR_OK = 0,
  PYR_ERR_INVALID_ARG = -1,
  PYR_ERR_HIP = -2,
  PYR_ERR_NO_BEST_NORMAL = -3, /* particle_state.hpp:176-189 */
  PYR_ERR_NOT_FOUND = -4,
  PYR_ERR_INTERNAL = -5
};

/* precision forms */
enum {
  PYR_FP64 = 0,   /* plain FP64 throughout: the drop-in */
  PYR_FP32C = 1   /* FP32 bulk arithmetic, FP32 partial sums folded into FP64
                     per-node sums, FP64 positions, face tests and charge
                     normalization */
};

typedef struct pyr_mesh pyr_mesh;

/* Mesh contract, all HOST pointers (layout: pyrticle_amd/meshdata.py).
 * Replaces _internal.MeshData + MeshData.fill_from_hedge
 * (wrap_meshdata.cpp:50-114, meshdata.py:37-133, meshdata.hpp:53-216). */
typedef struct {
  int32_t dim, K, Np, nv, nadj;
  const double *inv_map;        /* K x (d*d + d): A row-major, then b        */
  const double *jacobian;       /* K                                          */
  const uint32_t *el_vertices;  /* K x (d+1)                                  */
  const double *face_normal;    /* K x (d+1) x d                              */
  const double *face_rhs;       /* K x (d+1)                                  */
  const double *face_centroid;  /* K x (d+1) x d                              */
  const double *face_radius;    /* K x (d+1)                                  */
  const uint32_t *face_neighbor;/* K x (d+1), PYR_INVALID_ELEMENT on boundary  */
  const uint32_t *face_per_axis;/* K x (d+1), 0xFFFFFFFF if not periodic      */
  const double *vertices;       /* nv x d                                     */
  const uint32_t *vadj_starts;  /* nv + 1                                     */
  const uint32_t *vadj_elements;/* nadj                                       */
  const double *periodicity;    /* d x 2 (min, max); min == max: none         */
  const double *nodes;          /* K*Np x d                                   */
  const int32_t *exponents;     /* Np x d monomial exponents (node tuples)    */
  const double *vandermonde;    /* Np x Np row-major, V[i][k] = phi_k(node_i) */
} pyr_mesh_desc;

/* library ------------------------------------------------------------------ */
int pyr_init(int device);                 /* select the HIP device           */
const char *pyr_last_error(void);
int pyr_set_stream(void *hip_stream);     /* hipStream_t; NULL = null stream */
int pyr_sync(void);
int pyr_version(void);

/* Region timing with HIP events on the library stream (no reference
 * counterpart; the reference times with pytools IntervalTimers, pusher.py:41-43,
 * deposition/__init__.py:50-75).  Regions: 0 the fused push kernel, 1 the
 * deposition gather kernel, 2 the element sort and staging, 3 compaction,
 * 4 the push adjoint, 5 the deposition adjoint.
 * Enabling clears earlier records; read synchronizes the device. */
int pyr_profile_enable(int on);
int pyr_profile_read(int region, double *total_ms, int64_t *launches);

/* mesh --------------------------------------------------------------------- */
int pyr_mesh_create(const pyr_mesh_desc *desc, pyr_mesh **out);
int pyr_mesh_destroy(pyr_mesh *mesh);

/* meshdata.hpp:208-215 find_containing_element (wrap_meshdata.cpp:110), for n
 * points: the lowest-numbered element containing each point, or
 * PYR_INVALID_ELEMENT. */
int pyr_find_containing_element(pyr_mesh *mesh, int64_t n, const double *x,
                                uint32_t *elem_out);

/* meshdata.hpp:202-206 is_in_element (wrap_meshdata.cpp:108), for n
 * (element, point) pairs; inside_out[i] is 0 or 1. */
int pyr_is_in_element(pyr_mesh *mesh, int64_t n, const uint32_t *elem, const double *x,
                      uint8_t *inside_out);

/* particles ----------------------------------------------------------------- */
/* particle_state.hpp:98-125 get_velocities (wrap_pic.cpp:71). */
int pyr_velocities(int64_t n, int vdim, double vacuum_c, const double *momenta,
                   const double *masses, double *v_out);

/* particle_state.hpp:130-357 find_new_containing_element, for all n particles
 * as update_containing_elements does (wrap_pic.cpp:72-75), including the
 * periodic wrap of boundary_hit, which updates positions in place.  Instead of
 * calling the boundary-hit listener, particles that must die get dead_out = 1;
 * pyr_compact_plan turns those flags into the reference's kill order.
 * counters (device, 4 x uint64: same, neighbor, vertex, global) accumulate
 * (FindEventCounters, wrap_pic.cpp:90-96); may be NULL. */
int pyr_locate(pyr_mesh *mesh, int64_t n, double *positions, const double *momenta,
               uint32_t *containing_elements, uint8_t *dead_out,
               unsigned long long *counters);

/* The net effect of kill_particle/move_particle (particle_state.hpp:361-407)
 * applied in update_containing_elements' order to the particles flagged dead:
 * holes below the new count, in ascending order, receive the surviving tail
 * particles in descending order.  Writes n_moves (from, to) pairs to device
 * arrays of capacity n and returns the new count and n_moves on the host. */
int pyr_compact_plan(int64_t n, const uint8_t *dead, uint32_t *move_from,
                     uint32_t *move_to, int64_t *n_new_out, int64_t *n_moves_out);

/* Applies moves to one per-particle array of elem_bytes bytes per particle
 * (NumberShiftableVector.note_move, pyrticle/tools.py:174-175). */
int pyr_apply_moves(void *buffer, int64_t elem_bytes, int64_t n_moves,
                    const uint32_t *move_from, const uint32_t *move_to);

/* The same plan as pyr_compact_plan for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  It allocates nothing:
 * its scratch (the scan of the dead flags, the holes, the scan's tile
 * partials) is a static array of about 8 MB in the library image on each
 * device, held until the process ends.  n is limited to PYR_STATIC_PLAN_MAX;
 * larger n returns PYR_ERR_INVALID_ARG.  Calls on one device share that
 * array, so calls that overlap on one device race (serialize them, as one
 * stream does). */
#define PYR_STATIC_PLAN_MAX 1048576
int pyr_compact_plan_static(int64_t n, const uint8_t *dead, uint32_t *move_from,
                            uint32_t *move_to, int64_t *n_new_out, int64_t *n_moves_out);

/* Frees the library-level scratch that pyr_compact_plan grows on demand (the
 * per-mesh scratch goes with pyr_mesh_destroy).  No counterpart in the
 * reference, whose kill loop allocates nothing on a device; call it when no
 * other library call is in flight. */
int pyr_release_scratch(void);

/* push_monomial.hpp:219-325 monomial_particle_pusher::forces
 * (wrap_pusher.cpp:101-119) with make_interpolator (:155-210): interpolates
 * the six nodal fields (device, K*Np each; NULL = ZeroVector) at the particle
 * positions and returns F = q e + v x (q b) truncated to vdim.  vis_* (n x 3,
 * may be NULL) receive pt_e, pt_b, el_force, mag_force as the reference's
 * visualization listener does (:307-322). */
int pyr_forces(pyr_mesh *mesh, int64_t n, int vdim, const double *const fields[6],
               const double *positions, const uint32_t *containing_elements,
               const double *charges, const double *velocities, double *force_out,
               double *vis_e, double *vis_b, double *vis_el_force, double *vis_mag_force);

/* Stores six nodal fields (device, K*Np each; NULL = zero) as per-element
 * monomial coefficients inside the mesh handle, for pyr_lsrk4_steps.  The
 * static fields of a run are transformed once.  No reference counterpart: the
 * reference re-solves the interpolation system per call (push_monomial.hpp:
 * 192-203); mathematically the same interpolant. */
int pyr_set_static_fields(pyr_mesh *mesh, const double *const fields[6]);

/* Fused hedge LSRK4 (5 stages) of the particle RHS [v, F] in the static fields
 * of pyr_set_static_fields, nsteps steps of size dt, with relocalization and
 * kills after every stage exactly as cloud.py:554-603 advance_state and
 * update_containing_elements order them (kills compact the arrays).  One
 * kernel pass per step keeps each particle's 5 stages in registers.
 * *n_inout (host) is the particle count before and after.  ids (device int64,
 * may be NULL) is moved along with the particles.  precision: PYR_FP64 or
 * PYR_FP32C (interpolation and force in FP32, positions and momenta FP64). */
int pyr_lsrk4_steps(pyr_mesh *mesh, int precision, int nsteps, double dt, double vacuum_c,
                    int64_t *n_inout, double *positions, double *momenta, double *charges,
                    double *masses, uint32_t *containing_elements, int64_t *ids,
                    unsigned long long *counters);

/* PYR_FP32C pyr_lsrk4_steps on this mesh: evaluate the interpolation
 * contraction (6 fields x Np monomials) of full waves on the matrix cores
 * instead of the vector ALUs, one pass per element present in a block of four
 * consecutive particles (one pass when particles are sorted by element).
 * mode: 0 off (default), 1 on, 2 on and count the particle-stages that take
 * the matrix-core path (pyr_push_mfma_waves returns the count since the last
 * call and resets it).  The setting belongs to the mesh handle.
 * No reference counterpart: a switch between two evaluations of the same FP32
 * interpolant (push_monomial.hpp:99-111), kept for measurement and tests. */
int pyr_push_set_mfma(pyr_mesh *mesh, int mode);
int pyr_push_mfma_waves(pyr_mesh *mesh, int64_t *out);

/* adjoint of the push ------------------------------------------------------ */
/* pyr_lsrk4_steps with the averaging pusher's force (push_average.hpp:338-443,
 * pyrticle/pusher.py:167-207 AverageParticlePusher, driven by hedge LSRK4 as
 * test/test_pyrticle.py:381-384 runs it): the static fields of
 * pyr_set_static_fields (nodal) averaged over each particle's shape of radius,
 * alpha, normalizer with integral_weights (device, Np doubles, M 1).  Per
 * stage: velocities, the shape-weighted gather of pyr_average_forces, the RK
 * update and relocalization; kills compact the arrays in the per-stage path's
 * order.  Arguments otherwise as pyr_lsrk4_steps (FP64 only). */
int pyr_lsrk4_steps_average(pyr_mesh *mesh, int nsteps, double dt, double vacuum_c,
                            double radius, double alpha, double normalizer,
                            const double *integral_weights, int64_t *n_inout, double *x,
                            double *p, double *q, double *mass, uint32_t *containing_elements,
                            int64_t *ids, unsigned long long *counters);

/* No reference counterpart (pyrticle has no adjoint); these differentiate the
 * pyr_lsrk4_steps dynamics for torch.autograd (pyrticle_amd/adjoint.py).
 *
 * pyr_lsrk4_forward: nsteps LSRK4 steps as pyr_lsrk4_steps computes them, but
 * without compaction: a particle that leaves the mesh keeps its slot with
 * death[i] = 5*step + stage (1-based; 0 = alive) and containing_elements[i] =
 * INVALID.  If ckpt_xp/ckpt_elem are given, the state at the start of every
 * ckpt_every-th step is stored there: ckpt_xp[(seg*n + i)*2*dim + k] holds x
 * (k < dim) then p, for seg = 0 .. ceil(nsteps/ckpt_every)-1.  The particles
 * are stepped in element order (a sort on the mesh's sort scratch, as
 * pyr_sort_by_element). */
int pyr_lsrk4_forward(pyr_mesh *mesh, int precision, int nsteps, int ckpt_every, double dt,
                      double vacuum_c, int64_t n, double *positions, double *momenta,
                      const double *charges, const double *masses,
                      uint32_t *containing_elements, uint32_t *death, double *ckpt_xp,
                      uint32_t *ckpt_elem, unsigned long long *counters);

/* Reverse pass of pyr_lsrk4_forward: on entry x_bar/p_bar (n x dim, device)
 * hold dL/dx and dL/dp of the final state, on return those of the initial
 * state.  Segments are recomputed from the checkpoints (library scratch of
 * ckpt_every*n*(16*dim+4) + 5*n*(16*dim+4) bytes).  Particles with death != 0
 * get zero gradients.  precision PYR_FP32C evaluates fields and field
 * gradients in FP32 and accumulates the adjoint in FP64.  The particles are
 * processed in the element order of the first checkpoint. */
int pyr_lsrk4_adjoint(pyr_mesh *mesh, int precision, int nsteps, int ckpt_every, double dt,
                      double vacuum_c, int64_t n, const double *charges, const double *masses,
                      const uint32_t *death, const double *ckpt_xp, const uint32_t *ckpt_elem,
                      double *x_bar, double *p_bar);

/* Adjoint of pyr_velocities: p_bar = (dv/dp)^T v_bar (overwritten). */
int pyr_velocities_backward(int64_t n, int vdim, double vacuum_c, const double *momenta,
                            const double *masses, const double *v_bar, double *p_bar);

/* deposition ---------------------------------------------------------------- */
/* dep_target.hpp:203-273 deposit_densities / deposit_j / deposit_rho
 * (wrap_depositor.cpp:45-50, over slice(None)) for InterpolatingDepositor
 * (shape_function_depositor, dep_shape.hpp:67-150) with the polynomial shape
 * S(r) = normalizer * (R - r^2/R)^alpha for r <= R (tools.hpp:514-556;
 * normalizer as tools.cpp:43-58 computes it).  rho (K*Np) and j (K*Np x
 * vdim) are overwritten; either may be NULL; velocities may be NULL when j is.
 * Particles whose containing element is invalid contribute nothing.
 * precision PYR_FP32C evaluates the shape and its products in FP32 (on the
 * matrix cores) and folds short FP32 partial sums into FP64 per-node sums;
 * positions relative to the element, q*normalizer and the outputs stay FP64.
 * PYR_FP64 evaluates and sums in FP64 (on the FP64 matrix cores).  Both
 * precisions give bitwise reproducible results. */
int pyr_deposit(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                int precision, int64_t n, int vdim, const double *positions,
                const double *charges, const double *velocities,
                const uint32_t *containing_elements, double *rho_out, double *j_out);

/* dep_normshape.hpp:53-245 normalized_shape_function_depositor (wrap_depositor.cpp
 * :129-134, NormalizingInterpolatingDepositor; pyrticle/deposition/shape.py:58-98):
 * like pyr_deposit, but every particle deposits (q / I) S, I = sum over the
 * elements its shape reaches of |J_e| sum_i S(node_i - x) w_i with w = M 1
 * (integral_weights: device, Np doubles, the reference element's mass matrix
 * times ones), so that the deposited charge integrates to q exactly.  I is
 * formed in FP64 for both precisions.  Located particles with I = 0 deposit
 * nothing; their number is returned in *n_zero_out (host) for the caller's
 * warning (the reference's WARN per particle).  scale_out (device, n doubles,
 * may be NULL) receives q / I. */
int pyr_deposit_normalized(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                           int precision, int64_t n, int vdim, const double *x, const double *q,
                           const double *v, const uint32_t *containing_elements,
                           const double *integral_weights, double *rho, double *j,
                           double *scale_out, int64_t *n_zero_out);

/* push_average.hpp:338-443 averaging_particle_pusher::forces (wrap_pusher.cpp
 * :39-76, :124-140 AveragingPusher; pyrticle/pusher.py:167-207
 * AverageParticlePusher):
 * the force on each particle from the six nodal fields (device, K*Np each;
 * NULL = ZeroVector) averaged with the particle's deposited shape as weight,
 * <F> = sum_e |J_e| sum_i F_i S_i w_i / sum_e |J_e| sum_i S_i w_i over the nodes
 * the shape reaches (w = M 1: integral_weights, device, Np doubles), giving
 * F = q <E> + v x (q <B>) truncated to vdim (force_out, device, n x vdim).
 * The same force results with the plain or the normalized shape depositor
 * (the deposit's scale cancels in the ratio).  integral_out (device, n, may be
 * NULL) receives the shape integral I = N sum_e |J_e| sum_i S_i w_i, from
 * which the caller forms the reference's normalization statistics.
 * Particles with I = 0 get no force; their count goes to *n_zero_out (host).
 * vis_* (all or none, device): pt_e, pt_b (n x 3, the averages), pt_e_stddev,
 * pt_b_stddev (n), el_force, mag_force (n x vdim), as push_average.hpp
 * :430-438 stores them. */
int pyr_average_forces(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                       int64_t n, int vdim, const double *const fields[6], const double *x,
                       const uint32_t *containing_elements, const double *charges,
                       const double *velocities, const double *integral_weights,
                       double *force_out, double *integral_out, double *vis_e, double *vis_b,
                       double *vis_e_stddev, double *vis_b_stddev, double *vis_el_force,
                       double *vis_mag_force, int64_t *n_zero_out);

/* Particle-state diagnostics (diagnostics.hpp:37-183, wrap_pic.cpp:38-47).
 * Device inputs (positions n x xdim, momenta n x vdim, charges, masses,
 * velocities n x vdim); scalar and vector results go to host memory.  Sums are
 * formed on the device in a fixed order (deterministic; the reference sums
 * sequentially).  An empty state gives 0 (the reference's early returns). */

/* diagnostics.hpp:37-57 kinetic_energies: out (device, n) = (|p| / v - m) c^2,
 * v = c |p| / sqrt(m^2 c^2 + |p|^2). */
int pyr_kinetic_energies(int64_t n, int vdim, double vacuum_c, const double *momenta,
                         const double *masses, double *out);

/* diagnostics.hpp:61-70 total_charge: *out = sum q. */
int pyr_total_charge(int64_t n, const double *charges, double *out);

/* diagnostics.hpp:73-87 particle_momentum: out[vdim] = sum p. */
int pyr_particle_momentum(int64_t n, int vdim, const double *momenta, double *out);

/* diagnostics.hpp:171-183 particle_current: out[vdim] = sum q v / length. */
int pyr_particle_current(int64_t n, int vdim, const double *charges, const double *velocities,
                         double length, double *out);

/* diagnostics.hpp:90-101 rms_beam_size: *out = sqrt(mean x_axis^2). */
int pyr_rms_beam_size(int64_t n, int xdim, const double *positions, int axis, double *out);

/* diagnostics.hpp:106-155 rms_beam_emittance along axis with x' = p_axis /
 * p_beam_axis (0 where p_beam_axis = 0), by the reference's moment expansion. */
int pyr_rms_beam_emittance(int64_t n, int xdim, int vdim, const double *positions,
                           const double *momenta, int axis, int beam_axis, double *out);

/* diagnostics.hpp:159-167 rms_energy_spread with tools.hpp:303-322 std_dev:
 * *out = sqrt(<E^2> - <E>^2) over the kinetic energies. */
int pyr_rms_energy_spread(int64_t n, int vdim, double vacuum_c, const double *momenta,
                          const double *masses, double *out);

/* The reference's face walk (element_finder.hpp:203-294 face_based_element_finder,
 * as dep_normshape.hpp:131-136 and :175-177 gather statistics along it), for
 * the normalized depositor's centroid_distance_stats and el_per_particle_stats:
 * per particle (device arrays of n), the number of elements the walk visits
 * (n_elements) and the sum, sum of squares, minimum and maximum of the
 * distance from each visited element's vertex centroid to the shape centre
 * (the periodic image the walk uses).  Deposition itself does not walk; this
 * entry point exists for those statistics.  Fails if a walk visits more than
 * 384 elements. */
int pyr_face_walk_stats(pyr_mesh *mesh, double radius, int64_t n, const double *x,
                        const uint32_t *containing_elements, uint32_t *n_elements,
                        double *dist_sum, double *dist_sumsq, double *dist_min,
                        double *dist_max);

/* Adjoint of pyr_deposit (no reference counterpart): given rho_bar = dL/drho
 * (K*Np, may be NULL) and j_bar = dL/dj (K*Np x vdim, may be NULL), writes
 * x_bar = dL/dx (n x vdim) and, with j_bar, v_bar = dL/dv (n x vdim); both are
 * overwritten, and zero for particles with an invalid element.  A gather per
 * particle over the elements its shape reaches; no atomics.  alpha >= 1. */
int pyr_deposit_backward(pyr_mesh *mesh, double radius, double alpha, double normalizer,
                         int precision, int64_t n, int vdim, const double *positions,
                         const double *charges, const double *velocities,
                         const uint32_t *containing_elements, const double *rho_bar,
                         const double *j_bar, double *x_bar, double *v_bar);

/* Instrumentation (no reference counterpart): when on, pyr_deposit on this
 * mesh also counts [node-particle pairs evaluated, pairs inside R, candidate
 * particle visits, visits passing the per-element test]; pyr_deposit_stats
 * reads the counts of the last call.  FP64 precision only. */
int pyr_deposit_set_counting(pyr_mesh *mesh, int on);
int pyr_deposit_stats(pyr_mesh *mesh, unsigned long long out[4]);

/* Instrumentation (no reference counterpart): statistics of the culled chunk
 * lists of the last PYR_FP32C pyr_deposit on this mesh: [list entries
 * (chunk-target pairs), listed (chunk, node group) pairs, 4x4x1 matrix-core
 * instructions the list gather issued, work items]; zeros when the last
 * deposition was not PYR_FP32C.  Synchronizes the library stream. */
int pyr_deposit_list_stats(pyr_mesh *mesh, unsigned long long out[4]);

/* Number of (target element, source element) pairs in the gather candidate
 * lists for radius R (built on first use, cached per radius). */
int pyr_deposit_candidates(pyr_mesh *mesh, double radius, int64_t *n_pairs_out);

/* Stable sort of particle indices by containing element (the radix sort the
 * deposition uses): perm_out[j] is the particle at sorted position j
 * (numpy.argsort(kind="stable") of the element ids); seg_out (K+1, may be NULL)
 * the start of each element's run.  No reference counterpart. */
int pyr_sort_by_element(pyr_mesh *mesh, int64_t n, const uint32_t *containing_elements,
                        uint32_t *perm_out, uint32_t *seg_out);

#ifdef __cplusplus
}
#endif

#endif
