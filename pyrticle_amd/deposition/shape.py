"""Shape-function depositor (pyrticle/deposition/shape.py:31-53)."""

from __future__ import annotations

from .. import _internal
from .. import device as dv
from . import Depositor


class ShapeFunctionDepositor(Depositor):
    """Backed by InterpolatingDepositor{2d2v,3d3v} on the GPU.

    ``precision`` selects the plain FP64 form (the drop-in, default) or the
    FP32-compensated form ("fp32c")."""

    name = "Shape"

    def __init__(self, precision="fp64"):
        Depositor.__init__(self)
        self.precision = dv.FP64 if precision == "fp64" else dv.FP32C

    def initialize(self, method):
        Depositor.initialize(self, method)
        backend_class = _internal.class_for("InterpolatingDepositor",
                                            method.get_dimensionality_suffix())
        self.backend = backend_class(method.mesh_data)
        self.backend.precision = self.precision

    def make_state(self, state):
        return self.backend.DepositorState()

    def set_shape_function(self, state, sf):
        Depositor.set_shape_function(self, state, sf)
        self.backend.shape_function = sf

    def note_move(self, state, orig, dest, size):
        pass

    def note_change_size(self, state, count):
        pass


class NormalizedShapeFunctionDepositor(Depositor):
    """pyrticle/deposition/shape.py:55-98: every particle's deposit is scaled so
    that its charge density integrates exactly to its charge over the elements
    it reaches (dep_normshape.hpp:122-195).  Backed by
    NormalizingInterpolatingDepositor{2d2v,3d3v}: per-particle shape integrals
    in FP64 on the GPU, then the same gather deposition as ShapeFunctionDepositor
    with the scaled charges.  The depositor state's
    ``stats.normalization_stats`` gathers q / integral of every deposited
    particle; ``centroid_distance_stats`` and ``el_per_particle_stats`` follow
    the reference's face walk, run as a pass of its own when ``walk_stats`` is
    set (and from add_instrumentation on).  That pass is instrumentation and
    costs far more than the deposition: a depth-first walk per particle with
    its visited set in private memory, about 216 ms for 4M particles on the
    200,910-tet cylinder against 17 ms for the normalized deposition."""

    name = "NormShape"

    def __init__(self, precision="fp64", walk_stats=False):
        Depositor.__init__(self)
        self.precision = dv.FP64 if precision == "fp64" else dv.FP32C
        self.walk_stats = walk_stats

    def initialize(self, method):
        Depositor.initialize(self, method)
        eg, = method.discretization.element_groups
        ldis = eg.local_discretization
        backend_class = _internal.class_for("NormalizingInterpolatingDepositor",
                                            method.get_dimensionality_suffix())
        self.backend = backend_class(method.mesh_data, ldis.mass_matrix())
        self.backend.precision = self.precision
        self.backend.walk_stats = self.walk_stats

    def add_instrumentation(self, mgr, observer):
        """shape.py:71-88 logs all three statistics: the face-walk ones
        (centroid distances, elements per particle) are gathered from here on."""
        Depositor.add_instrumentation(self, mgr, observer)
        self.walk_stats = self.backend.walk_stats = True

    def make_state(self, state):
        return self.backend.DepositorState()

    def set_shape_function(self, state, sf):
        Depositor.set_shape_function(self, state, sf)
        self.backend.shape_function = sf

    def note_move(self, state, orig, dest, size):
        pass

    def note_change_size(self, state, count):
        pass
