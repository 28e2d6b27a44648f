"""Particle pushers (pyrticle/pusher.py:31-114)."""

from __future__ import annotations

import time

import numpy as np

from . import _internal


class EmptyPusherState:
    pass


class Pusher:
    def __init__(self):
        self.force_time = 0.0     # t_force (pusher.py:41-43)

    def initialize(self, method):
        self.method = method

    def make_state(self, state):
        return EmptyPusherState()

    def advance_state(self, state):
        return state.pusher_state

    def add_instrumentation(self, mgr, observer):
        pass

    def upkeep(self, state):
        pass

    def _forces(self, state, velocities, *field_args):
        return self.backend.forces(*field_args, ps=state.particle_state,
                                   velocities=velocities, vis_listener=state.vis_listener)

    def forces(self, state, velocities, *field_args):
        t0 = time.perf_counter()
        f = self._forces(state, velocities, *field_args)
        self.force_time += time.perf_counter() - t0
        return f

    def note_move(self, state, orig, dest, size):
        pass

    def note_change_size(self, state, count):
        pass


class MonomialParticlePusher(Pusher):
    """pusher.py:86-114: monomials over the node tuples, LU of V^T."""

    def initialize(self, method):
        Pusher.initialize(self, method)
        backend_class = _internal.class_for("MonomialPusher",
                                            method.get_dimensionality_suffix())
        self.backend = backend_class(method.mesh_data)
        discr = method.mesh_data.discr
        for i, eg in enumerate(discr.element_groups):
            ldis = eg.local_discretization
            lmd = _internal.LocalMonomialDiscretization()
            lmd.basis.extend([_internal.MonomialBasisFunction(*idx)
                              for idx in ldis.node_tuples()])
            vdm = np.array([[b(x) for b in lmd.basis] for x in ldis.unit_nodes()])
            lmd.lu_vandermonde_t, lmd.lu_piv_vandermonde_t = _internal.lu(vdm.T)
            self.backend.local_discretizations.append(lmd)
            self.backend.ldis_indices.extend([i] * len(eg.members))


class AverageParticlePusher(Pusher):
    """pusher.py:167-207: forces from the fields averaged over each particle's
    deposited shape (AveragingPusher{2d2v,3d3v}, one shape-weighted gather on
    the GPU per call).  Uses the method's depositor for the shape function and
    must run with a shape depositor (plain or normalized)."""

    def initialize(self, method):
        Pusher.initialize(self, method)
        eg, = method.mesh_data.discr.element_groups
        ldis = eg.local_discretization
        backend_class = _internal.class_for("AveragingPusher",
                                            method.get_dimensionality_suffix())
        self.backend = backend_class(method.mesh_data, ldis.mass_matrix())

    def make_state(self, state):
        return self.backend.PusherState()

    def _forces(self, state, velocities, *field_args):
        state.pusher_state.e_normalization_stats.reset()
        state.pusher_state.b_normalization_stats.reset()
        return self.backend.forces(
            *field_args, particle_state=state.particle_state, pusher_state=state.pusher_state,
            depositor=self.method.depositor.backend, depositor_state=state.depositor_state,
            velocities=velocities, vis_listener=state.vis_listener)
