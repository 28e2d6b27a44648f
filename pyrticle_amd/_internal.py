"""The names the reference's Python layer imports from ``pyrticle._internal``
(src/wrapper/wrap_*.cpp), implemented over libpyrticle_mi355x.so.

Particle arrays are torch CUDA tensors (positions/momenta n x d float64,
charges/masses float64, containing elements int32 holding uint32 bits); every
hot-path call goes through the C ABI.  Differences from the Boost.Python
module, all deliberate:

* ``update_containing_elements`` reports kills as data: it compacts the state
  in the reference's order itself and notifies the number-shift listener with
  one batched ``note_move`` (index tensors) and ``note_change_size`` instead of
  calling a boundary-hit listener per particle (particle_state.hpp:266-357,
  cloud.py:581-585);
* ``deposit_*`` accept ``slice(None)`` or a contiguous ``slice(start, stop)``;
* ``lu`` and the monomial basis objects keep their reference meaning for
  callers, but the GPU interpolation uses the mesh's Vandermonde matrix
  directly (same interpolant).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import device as dv
from .discretization import INVALID_AXIS as _INVALID_AXIS
from .meshdata import build_mesh_arrays

INVALID_ELEMENT = dv.INVALID_ELEMENT


# ---------------------------------------------------------------------------
# listeners (wrap_tools.cpp:45-120, 220-262; tools.hpp:422-508)
# ---------------------------------------------------------------------------
class NumberShiftListener:
    def note_change_size(self, new_size):
        pass

    def note_move(self, orig, dest, size):
        pass

    def note_reset(self, start, size):
        pass


class BoundaryHitListener:
    def note_boundary_hit(self, pn):
        pass


class VisualizationListener:
    def store_mesh_vis_vector(self, name, vec):
        pass

    def store_particle_vis_vector(self, name, vec):
        pass


class WarningListener:
    _singleton = None

    def __init__(self):
        WarningListener._singleton = self

    def note_warning(self, message, filename, lineno):
        pass

    @staticmethod
    def warn(message, filename="<pyrticle_amd>", lineno=0):
        if WarningListener._singleton is None:
            raise RuntimeError("warning with no listener: %s" % message)
        WarningListener._singleton.note_warning(message, filename, lineno)


class ZeroVector:
    """Stands for an all-zero field component (push_monomial.hpp:128-132)."""


class FindEventCounters:
    """particle_state.hpp:86-92; counters live on the device."""

    def __init__(self):
        self._t = torch.zeros(4, dtype=torch.int64, device="cuda")

    @property
    def tensor(self):
        return self._t

    find_same = property(lambda self: int(self._t[0]))
    find_by_neighbor = property(lambda self: int(self._t[1]))
    find_by_vertex = property(lambda self: int(self._t[2]))
    find_global = property(lambda self: int(self._t[3]))


# ---------------------------------------------------------------------------
# mesh data (wrap_meshdata.cpp:50-114, meshdata.py:37-152)
# ---------------------------------------------------------------------------
class PeriodicityAxis:
    def __init__(self, min=0.0, max=0.0):
        self.min = min
        self.max = max


class MeshData:
    INVALID_ELEMENT = INVALID_ELEMENT
    INVALID_AXIS = _INVALID_AXIS

    def __init__(self, dimensions):
        self.dimensions = dimensions
        self.discr = None
        self.arrays = None
        self.device = None
        self.periodicities = []

    @staticmethod
    def get_INVALID_ELEMENT():
        return INVALID_ELEMENT

    def fill_from_hedge(self, discr):
        """meshdata.py:37-133, against the discretization shim."""
        self.discr = discr
        self.arrays = build_mesh_arrays(discr)
        self.periodicities = [PeriodicityAxis(*row) for row in self.arrays.periodicity]
        self.device = dv.DeviceMesh(self.arrays)

    def node_count(self):
        return self.arrays.nnodes

    def find_containing_element(self, pt):
        x = torch.as_tensor(np.asarray(pt, dtype=np.float64).reshape(1, -1), device="cuda")
        return int(dv.element_numpy(self.device.find_containing_element(x))[0])

    def find_containing_elements(self, x: torch.Tensor) -> torch.Tensor:
        return self.device.find_containing_element(x)

    def is_in_element(self, en, pt):
        x = torch.as_tensor(np.asarray(pt, dtype=np.float64).reshape(1, -1), device="cuda")
        e = dv.element_tensor(np.array([en], dtype=np.uint32), "cuda")
        return bool(self.device.is_in_element(e, x)[0])

    def min_vertex_distance_for_el(self, en):
        v = self.arrays.vertices[self.arrays.el_vertices[en]]
        d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=2)
        return float(d[~np.eye(len(v), dtype=bool)].min())

    def _min_vertex_distances(self):
        v = self.arrays.vertices[self.arrays.el_vertices]
        d = np.linalg.norm(v[:, :, None, :] - v[:, None, :, :], axis=3)
        n = v.shape[1]
        d[:, np.arange(n), np.arange(n)] = np.inf
        return d.reshape(len(v), -1).min(axis=1)

    def advisable_particle_radius(self):
        """meshdata.py:144-148."""
        dists = np.sort(self._min_vertex_distances())
        return 0.6 * float(dists[int(0.25 * len(dists))])

    def min_vertex_distance(self):
        return float(self._min_vertex_distances().min())


# ---------------------------------------------------------------------------
# particle state (particle_state.hpp:42-81, wrap_pic.cpp:53-69)
# ---------------------------------------------------------------------------
class _ParticleState:
    xdim = 0
    vdim = 0

    def __init__(self):
        self.particle_count = 0
        dev = "cuda"
        self.containing_elements = torch.zeros(0, dtype=torch.int32, device=dev)
        self.positions = torch.zeros((0, self.xdim), dtype=torch.float64, device=dev)
        self.momenta = torch.zeros((0, self.vdim), dtype=torch.float64, device=dev)
        self.charges = torch.zeros(0, dtype=torch.float64, device=dev)
        self.masses = torch.zeros(0, dtype=torch.float64, device=dev)


class ParticleState2d2v(_ParticleState):
    xdim = 2
    vdim = 2


class ParticleState3d3v(_ParticleState):
    xdim = 3
    vdim = 3


def get_velocities(ps, vacuum_c):
    """particle_state.hpp:98-125."""
    n = ps.particle_count
    return dv.velocities(ps.momenta[:n].contiguous(), ps.masses[:n].contiguous(), vacuum_c)


def _state_arrays(ps):
    return (ps.positions, ps.momenta, ps.charges, ps.masses, ps.containing_elements)


def update_containing_elements(mesh_data, ps, nshift_listener=None, counters=None):
    """particle_state.hpp:266-357 with the kill listener of cloud.py:581-585.

    Relocalizes all particles, wraps periodic ones, and removes the particles
    that left the mesh in the reference's swap-remove order.  The listener
    receives the batched moves (from, to) and the new count.  Returns the new
    particle count.
    """
    n = ps.particle_count
    if n == 0:
        return 0
    x = ps.positions[:n]
    dead = mesh_data.device.locate(x, ps.momenta[:n], ps.containing_elements[:n],
                                   counters.tensor if counters is not None else None)
    n_new, frm, to = dv.compact_plan(dead)
    if frm.numel():
        for arr in _state_arrays(ps):
            dv.apply_moves(arr, frm, to)
        if nshift_listener is not None:
            nshift_listener.note_move(frm, to, 1)
    if n_new != n:
        ps.particle_count = n_new
        if nshift_listener is not None:
            nshift_listener.note_change_size(n_new)
    return n_new


def move_particle(ps, frm, to, nshift_listener=None):
    """particle_state.hpp:386-407."""
    f = torch.tensor([frm], dtype=torch.int32, device="cuda")
    t = torch.tensor([to], dtype=torch.int32, device="cuda")
    for arr in _state_arrays(ps):
        dv.apply_moves(arr, f, t)
    if nshift_listener is not None:
        nshift_listener.note_move(f, t, 1)


def kill_particle(ps, pn, nshift_listener=None):
    """particle_state.hpp:361-377."""
    ps.particle_count -= 1
    if ps.particle_count:
        move_particle(ps, ps.particle_count, pn, nshift_listener)
    if nshift_listener is not None:
        nshift_listener.note_change_size(ps.particle_count)


def find_new_containing_element(mesh_data, ps, pn, prev, counters=None):
    """particle_state.hpp:130-261 for one particle (no kill)."""
    x = ps.positions[pn:pn + 1].clone()
    e = ps.containing_elements[pn:pn + 1].clone()
    e[0] = torch.tensor(np.array([prev], dtype=np.uint32).view(np.int32)[0])
    dead = mesh_data.device.locate(x, ps.momenta[pn:pn + 1].contiguous(), e,
                                   counters.tensor if counters is not None else None)
    return INVALID_ELEMENT if int(dead[0]) else int(dv.element_numpy(e)[0])


# ---------------------------------------------------------------------------
# shape function (tools.hpp:514-556, tools.cpp:35-58, wrap_tools.cpp:282-295)
# ---------------------------------------------------------------------------
class PolynomialShapeFunction:
    def __init__(self, radius=1.0, dimensions=1, alpha=2.0):
        self.radius = float(radius)
        self.dimensions = int(dimensions)
        self.exponent = float(alpha)
        n = float(dimensions)
        sphere_area = 2 * math.pi ** (n / 2) / math.gamma(n / 2)
        beta = math.gamma(n / 2) * math.gamma(alpha + 1) / math.gamma(n / 2 + alpha + 1)
        self.normalizer = 2 / (sphere_area * radius ** (n + alpha) * beta)

    @staticmethod
    def name():
        return "polynomial"

    def __call__(self, r):
        r2 = float(np.dot(r, r))
        if r2 > self.radius ** 2:
            return 0.0
        t = self.radius - r2 / self.radius
        return self.normalizer * t ** self.exponent


def get_shape_function_name():
    return PolynomialShapeFunction.name()


# ---------------------------------------------------------------------------
# depositor (wrap_depositor.cpp:45-126, dep_target.hpp:203-273)
# ---------------------------------------------------------------------------
class _InterpolatingDepositor:
    class DepositorState:
        def note_move(self, orig, dest, size):
            pass

        def note_change_size(self, count):
            pass

        def clear(self):
            pass

    def __init__(self, mesh_data):
        self.mesh_data = mesh_data
        self.shape_function = None
        self.precision = dv.FP64


class InterpolatingDepositor2d2v(_InterpolatingDepositor):
    pass


class InterpolatingDepositor3d3v(_InterpolatingDepositor):
    pass


class StatsGatherer:
    """tools.hpp:340-400 stats_gatherer (running mean and variance, Welford);
    add_array merges a whole batch with the parallel combination formula."""

    def __init__(self):
        self.reset()

    def reset(self):
        self._count, self._mean, self._m2 = 0, 0.0, 0.0
        self._min = self._max = None

    def add(self, x):
        x = float(x)
        if self._count == 0 or x < self._min:
            self._min = x
        if self._count == 0 or x > self._max:
            self._max = x
        self._count += 1
        delta = x - self._mean
        self._mean += delta / self._count
        self._m2 += delta * (x - self._mean)

    @staticmethod
    def moments(a):
        """(count, mean, sum of squared deviations, min, max) of a numpy array
        or device tensor, reduced where it lives; None when empty."""
        if isinstance(a, torch.Tensor):
            a = a.reshape(-1).to(torch.float64)
            if a.numel() == 0:
                return None
            mean_t = a.mean()
            red = torch.stack([mean_t, ((a - mean_t) ** 2).sum(), a.min(), a.max()]).cpu()
            return (a.numel(),) + tuple(float(t) for t in red)
        a = np.asarray(a, dtype=np.float64).ravel()
        if a.size == 0:
            return None
        mb = float(a.mean())
        return a.size, mb, float(((a - mb) ** 2).sum()), float(a.min()), float(a.max())

    def add_moments(self, mom):
        """Merges a batch given by its moments (parallel variance formula)."""
        if mom is None:
            return
        nb, mb, m2b, lo, hi = mom
        if self._count == 0:
            self._min, self._max = lo, hi
        else:
            self._min, self._max = min(self._min, lo), max(self._max, hi)
        n = self._count + nb
        delta = mb - self._mean
        self._mean += delta * nb / n
        self._m2 += m2b + delta * delta * self._count * nb / n
        self._count = n

    def add_array(self, a):
        """Adds every entry of a (numpy array or device tensor)."""
        self.add_moments(self.moments(a))

    def count(self):
        return self._count

    def minimum(self):
        return self._min

    def maximum(self):
        return self._max

    def mean(self):
        return self._mean

    def variance_sample(self):
        return self._m2 / (self._count - 1)

    def standard_deviation_sample(self):
        return math.sqrt(self.variance_sample())

    def variance(self):
        return self._m2 / self._count

    def standard_deviation(self):
        return math.sqrt(self.variance())


class _NormalizingInterpolatingDepositor(_InterpolatingDepositor):
    """dep_normshape.hpp:53-245 (wrap_depositor.cpp:129-134): constructed from the
    local discretization's mass matrix; the integral weights are M 1
    (dep_normshape.hpp:214-222)."""

    class DepositorState(_InterpolatingDepositor.DepositorState):
        class _Stats:
            """dep_normshape.hpp:42-47 normalized_deposition_stats."""

            def __init__(self):
                self.normalization_stats = StatsGatherer()
                self.centroid_distance_stats = StatsGatherer()
                self.el_per_particle_stats = StatsGatherer()

        def __init__(self):
            self.stats = self._Stats()

        def clear(self):
            self.stats = self._Stats()

    def __init__(self, mesh_data, mass_matrix):
        _InterpolatingDepositor.__init__(self, mesh_data)
        mm = np.asarray(mass_matrix, dtype=np.float64)
        self.integral_weights = mm @ np.ones(mm.shape[1])
        # the face-walk statistics (centroid distances, elements per particle)
        # need a pass of their own; gathered when set
        self.walk_stats = False


class NormalizingInterpolatingDepositor2d2v(_NormalizingInterpolatingDepositor):
    pass


class NormalizingInterpolatingDepositor3d3v(_NormalizingInterpolatingDepositor):
    pass


def _walk_stats(md, radius, x, elem, stats):
    """dep_normshape.hpp:131-136, 175-177: the face walk's elements per particle
    and centroid distances, merged into the depositor state's statistics."""
    n_el, dsum, dsq, dmin, dmax = md.device.face_walk_stats(radius, x, elem)
    stats.el_per_particle_stats.add_array(n_el)
    visited = n_el > 0
    count = int(n_el.sum())
    if count == 0:
        return
    red = torch.stack([dsum.sum(), dsq.sum(), dmin[visited].min(), dmax[visited].max()]).cpu()
    total, total_sq, lo, hi = (float(t) for t in red)
    mean = total / count
    stats.centroid_distance_stats.add_moments((count, mean, max(total_sq - count * mean * mean,
                                                                0.0), lo, hi))


def _sliced(pslice, n, arrays):
    """The particles of a Python slice over [0, n) (tools.hpp:38-61 iterates
    start, start + step, ... as the slice's indices() give them): contiguous
    views for step 1, else the selected rows gathered into contiguous device
    buffers.  The deposited sums do not depend on the particle order."""
    start, stop, step = (0, n, 1) if pslice is None else pslice.indices(n)
    if step == 1:
        stop = max(start, stop)
        return [None if t is None else t[start:stop] for t in arrays]
    idx = torch.arange(start, stop, step, dtype=torch.int64, device=arrays[0].device)
    return [None if t is None else t.index_select(0, idx).contiguous() for t in arrays]


def _deposit(dep, ps, node_count, velocities, pslice, want_rho, want_j, ds=None):
    sf = dep.shape_function
    if sf is None:
        raise RuntimeError("shape function never set")
    md = dep.mesh_data
    if node_count != md.node_count():
        raise ValueError("node count does not match the mesh")
    x, q, el, v = _sliced(pslice, ps.particle_count,
                          [ps.positions, ps.charges, ps.containing_elements, velocities])
    if v is not None:
        v = v.contiguous()
    if isinstance(dep, _NormalizingInterpolatingDepositor):
        scale = torch.empty(x.shape[0], dtype=torch.float64, device=x.device)
        rho, j, n_zero = md.device.deposit_normalized(
            sf.radius, sf.exponent, sf.normalizer, x, q, el, dep.integral_weights, v=v,
            want_rho=want_rho, want_j=want_j, precision=dep.precision, scale=scale)
        if ds is not None:
            ds.stats.normalization_stats.add_array(scale[scale != 0])
            if dep.walk_stats:
                _walk_stats(md, sf.radius, x, el, ds.stats)
        if n_zero:
            # dep_normshape.hpp:175-183 warns once per such particle
            WarningListener.warn("deposited particle mass is zero (%d particles)" % n_zero)
        return rho, j
    rho, j = md.device.deposit(sf.radius, sf.exponent, sf.normalizer, x, q, el, v=v,
                               want_rho=want_rho, want_j=want_j, precision=dep.precision)
    return rho, j


def deposit_rho(dep, ds, ps, node_count, pslice=slice(None)):
    return _deposit(dep, ps, node_count, None, pslice, True, False, ds)[0]


def deposit_j(dep, ds, ps, node_count, velocities, pslice=slice(None)):
    """j as node_count x vdim (node-major, as the reference returns it)."""
    return _deposit(dep, ps, node_count, velocities, pslice, False, True, ds)[1]


def deposit_densities(dep, ds, ps, node_count, velocities, pslice=slice(None)):
    return _deposit(dep, ps, node_count, velocities, pslice, True, True, ds)


# ---------------------------------------------------------------------------
# pusher (wrap_pusher.cpp:85-171, push_monomial.hpp:45-325)
# ---------------------------------------------------------------------------
class MonomialBasisFunction:
    def __init__(self, *exponents):
        self.exponents = tuple(int(e) for e in exponents)

    def __call__(self, v):
        return float(np.prod([vi ** e for vi, e in zip(v, self.exponents)]))


class LocalMonomialDiscretization:
    def __init__(self):
        self.basis = []
        self.lu_vandermonde_t = None
        self.lu_piv_vandermonde_t = None


def lu(a):
    """wrap_tools.cpp:126-145 (getrf): (LU, pivots), column-major LAPACK form."""
    from scipy.linalg import lu_factor
    a = np.asarray(a, dtype=np.float64)
    lu_, piv = lu_factor(a, check_finite=True)
    if np.any(np.diag(lu_) == 0):
        raise RuntimeError("singular matrix in getrf")
    return np.asfortranarray(lu_), piv.astype(np.int32) + 1


class _MonomialPusher:
    def __init__(self, mesh_data):
        self.mesh_data = mesh_data
        self.local_discretizations = []
        self.ldis_indices = []

    def forces(self, ex, ey, ez, bx, by, bz, ps, velocities, vis_listener=None):
        if len(self.local_discretizations) > 1:
            raise RuntimeError("more than one local discretization is currently not "
                               "supported")
        n = ps.particle_count
        fields = [None if (f is None or isinstance(f, ZeroVector)) else f
                  for f in (ex, ey, ez, bx, by, bz)]
        res = self.mesh_data.device.forces(
            fields, ps.positions[:n], ps.containing_elements[:n], ps.charges[:n],
            velocities[:n].contiguous(), vis=vis_listener is not None)
        if vis_listener is not None:
            forces, vis = res
            for name in ("pt_e", "pt_b", "el_force", "mag_force"):
                vis_listener.store_particle_vis_vector(name, vis[name])
            return forces
        return res


class MonomialPusher2d2v(_MonomialPusher):
    pass


class MonomialPusher3d3v(_MonomialPusher):
    pass


class _AveragingPusher:
    """push_average.hpp:338-443 averaging_particle_pusher (wrap_pusher.cpp:124-140):
    forces from E and B averaged over each particle's deposited shape, with the
    depositor's shape function and element reach; integral weights M 1."""

    class PusherState:
        def __init__(self):
            self.e_normalization_stats = StatsGatherer()
            self.b_normalization_stats = StatsGatherer()

    def __init__(self, mesh_data, mass_matrix):
        self.mesh_data = mesh_data
        mm = np.asarray(mass_matrix, dtype=np.float64)
        self.integral_weights = mm @ np.ones(mm.shape[1])

    def forces(self, ex, ey, ez, bx, by, bz, particle_state, pusher_state, depositor,
               depositor_state, velocities, vis_listener=None):
        md = self.mesh_data
        sf = depositor.shape_function
        if sf is None:
            raise RuntimeError("shape function never set")
        fields = [None if (f is None or isinstance(f, ZeroVector)) else f
                  for f in (ex, ey, ez, bx, by, bz)]
        if md.dimensions == 2:
            # the 2D binding is the TE case: ez, bx, by are zero_vector
            # (wrap_pusher.cpp:57-73)
            fields[2] = fields[3] = fields[4] = None
        ps = particle_state
        n = ps.particle_count
        q = ps.charges[:n]
        res = md.device.average_forces(
            sf.radius, sf.exponent, sf.normalizer, fields, ps.positions[:n],
            ps.containing_elements[:n], q, velocities[:n].contiguous(), self.integral_weights,
            vis=vis_listener is not None)
        F, integral, n_zero = res[:3]
        live = integral != 0
        if isinstance(depositor, _NormalizingInterpolatingDepositor):
            # the normalized deposit runs underneath (dep_normshape.hpp:168-198):
            # its statistics and warnings, and charges q / I reaching the pusher
            dscale = q[live] / integral[live]
            if depositor_state is not None:
                depositor_state.stats.normalization_stats.add_array(dscale)
            if n_zero:
                WarningListener.warn("deposited particle mass is zero (%d particles)" % n_zero)
            pscale = q[live] / (dscale * integral[live])
        else:
            # reconstructed charge q I: scale q / (q I) (push_average.hpp:258-262)
            pscale = q[live] / (q[live] * integral[live])
            if n_zero and vis_listener is not None:
                WarningListener.warn("average pusher: %d particles had zero reconstructed "
                                     "charge" % n_zero)
        # both targets see the same reconstructed charge
        mom = StatsGatherer.moments(pscale)
        pusher_state.e_normalization_stats.add_moments(mom)
        pusher_state.b_normalization_stats.add_moments(mom)
        if vis_listener is not None:
            for name, arr in res[3].items():
                vis_listener.store_particle_vis_vector(name, arr)
        return F


class AveragingPusher2d2v(_AveragingPusher):
    pass


class AveragingPusher3d3v(_AveragingPusher):
    pass


# ---------------------------------------------------------------------------
# diagnostics (diagnostics.hpp:37-183, wrap_pic.cpp:38-47)
# ---------------------------------------------------------------------------
def _live(ps):
    n = ps.particle_count
    return n, ps.positions[:n], ps.momenta[:n]


def kinetic_energies(ps, vacuum_c):
    n, _, p = _live(ps)
    return dv.kinetic_energies(p, ps.masses[:n], vacuum_c)


def total_charge(ps):
    return dv.total_charge(ps.charges[:ps.particle_count])


def particle_momentum(ps):
    return dv.particle_momentum(_live(ps)[2])


def particle_current(ps, velocities, length):
    n = ps.particle_count
    return dv.particle_current(ps.charges[:n], velocities[:n].contiguous(), length)


def rms_beam_size(ps, axis):
    return dv.rms_beam_size(_live(ps)[1], axis)


def rms_beam_emittance(ps, axis, beam_axis):
    _, x, p = _live(ps)
    return dv.rms_beam_emittance(x, p, axis, beam_axis)


def rms_energy_spread(ps, vacuum_c):
    n, _, p = _live(ps)
    return dv.rms_energy_spread(p, ps.masses[:n], vacuum_c)


_SUFFIXES = {"2d2v": 2, "3d3v": 3}


def class_for(prefix, suffix):
    return globals()[prefix + suffix]
