"""Depositor plugin base (pyrticle/deposition/__init__.py:30-143)."""

from __future__ import annotations

import time

from .. import _internal


class _IntervalTimer:
    """Wall time and call counts of the deposition entry points (the reference
    uses pytools IntervalTimer / EventCounter, deposition/__init__.py:50-75)."""

    def __init__(self, name, description):
        self.name = name
        self.description = description
        self.elapsed = 0.0
        self.count = 0


class Depositor:
    def __init__(self):
        self.log_constants = {}
        self.deposit_timer = _IntervalTimer("t_deposit", "Time spent depositing")
        self.deposit_counter = 0

    def initialize(self, method):
        self.method = method
        self.shape_function = None

    def set_shape_function(self, state, sf):
        self.shape_function = sf

    def add_instrumentation(self, mgr, observer):
        pass

    def clear_particles(self):
        pass

    def deposit_hook(self):
        if self.shape_function is None:
            raise RuntimeError("shape function never set")

    def make_state(self, state):
        return self.backend.DepositorState()

    def note_move(self, state, orig, dest, size):
        state.depositor_state.note_move(orig, dest, size)

    def note_change_size(self, state, count):
        state.depositor_state.note_change_size(count)

    def _timed(self, fn, count, *args):
        t0 = time.perf_counter()
        out = fn(*args)
        self.deposit_timer.elapsed += time.perf_counter() - t0
        self.deposit_timer.count += 1
        self.deposit_counter += count
        return out

    def _deposit_densities(self, state, velocities, pslice):
        return _internal.deposit_densities(self.backend, state.depositor_state,
                                           state.particle_state, len(self.method.discretization),
                                           velocities, pslice)

    def _deposit_j(self, state, velocities, pslice):
        return _internal.deposit_j(self.backend, state.depositor_state, state.particle_state,
                                   len(self.method.discretization), velocities, pslice)

    def _deposit_rho(self, state, pslice):
        return _internal.deposit_rho(self.backend, state.depositor_state, state.particle_state,
                                     len(self.method.discretization), pslice)

    def deposit_densites(self, state, velocities):
        """(sic: deposition/__init__.py:117) rho and j, j as nodes x vdim."""
        self.deposit_hook()
        return self._timed(self._deposit_densities, 1 + self.method.dimensions_velocity,
                           state, velocities, slice(None))

    def deposit_j(self, state, velocities):
        self.deposit_hook()
        return self._timed(self._deposit_j, self.method.dimensions_velocity, state, velocities,
                           slice(None))

    def deposit_rho(self, state):
        self.deposit_hook()
        return self._timed(self._deposit_rho, 1, state, slice(None))

    def upkeep(self, state):
        pass

    def rhs(self, state):
        return 0

    def advance_state(self, state, rhs):
        return state.depositor_state
