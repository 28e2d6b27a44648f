"""Particle state container and PIC method (pyrticle/cloud.py:35-603).

``PicState`` holds device tensors; ``PicMethod`` is the reference's method
object with the same methods (velocities, add_particles, deposit_rho/j/
densities, advance_state, upkeep) and the RHS calculators.  The reference's
``deposit_densities`` calls ``self.velocities()`` without a state
(cloud.py:519) and its cache lookup tests the wrong name (cloud.py:144); both
are fixed here, the returned values are the ones the working paths give.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _internal
from . import device as dv
from .tools import NumberShiftableVector, is_zero, make_obj_array


class PicState:
    def __init__(self, method, particle_count=None, containing_elements=None,
                 positions=None, momenta=None, charges=None, masses=None,
                 depositor_state=None, pusher_state=None, pnss=None, vis_listener=None):
        state_class = _internal.class_for("ParticleState",
                                          method.get_dimensionality_suffix())
        pstate = self.particle_state = state_class()
        self.depositor_state = (method.depositor.make_state(self)
                                if depositor_state is None else depositor_state)
        self.pusher_state = (method.pusher.make_state(self)
                             if pusher_state is None else pusher_state)
        if particle_count is not None:
            pstate.particle_count = particle_count
            pstate.containing_elements = containing_elements
            pstate.positions = positions
            pstate.momenta = momenta
            pstate.charges = charges
            pstate.masses = masses
        self.derived_quantity_cache = {}
        if pnss is None:
            from .tools import StatePassingNumberShiftMultiplexer
            self.particle_number_shift_signaller = StatePassingNumberShiftMultiplexer(self)
        else:
            self.particle_number_shift_signaller = pnss
        if vis_listener is None:
            from .tools import MapStorageVisualizationListener
            self.vis_listener = MapStorageVisualizationListener(
                self.particle_number_shift_signaller)
        else:
            self.vis_listener = vis_listener

    def __len__(self):
        return self.particle_state.particle_count

    @property
    def positions(self):
        return self.particle_state.positions[:self.particle_state.particle_count]

    @property
    def momenta(self):
        return self.particle_state.momenta[:self.particle_state.particle_count]

    @property
    def masses(self):
        return self.particle_state.masses[:self.particle_state.particle_count]

    @property
    def charges(self):
        return self.particle_state.charges[:self.particle_state.particle_count]

    @property
    def containing_elements(self):
        return self.particle_state.containing_elements[:self.particle_state.particle_count]

    def resize(self, newsize):
        ps = self.particle_state

        def grow(t, shape):
            out = torch.zeros(shape, dtype=t.dtype, device=t.device)
            k = min(t.shape[0], shape[0])
            out[:k] = t[:k]
            return out

        ps.containing_elements = grow(ps.containing_elements, (newsize,))
        ps.positions = grow(ps.positions, (newsize, ps.xdim))
        ps.momenta = grow(ps.momenta, (newsize, ps.vdim))
        ps.charges = grow(ps.charges, (newsize,))
        ps.masses = grow(ps.masses, (newsize,))

    def clear(self):
        self.particle_state.particle_count = 0
        self.depositor_state.clear()
        self.derived_quantity_cache.clear()

    def get_derived_quantity_from_cache(self, name, getter):
        try:
            return self.derived_quantity_cache[name]
        except KeyError:
            self.derived_quantity_cache[name] = value = getter()
            return value


class TimesteppablePicState:
    """cloud.py:159-188: allows only ``+ update`` and ``* 1``."""

    def __init__(self, method, state):
        self.method = method
        self.state = state

    def __mul__(self, factor):
        if factor != 1:
            raise NotImplementedError("cannot multiply PIC state by anything but unity.")
        return self

    __rmul__ = __mul__

    def __add__(self, update):
        if is_zero(update):
            return self
        dx, dp, ddep = update
        dx = NumberShiftableVector.unwrap(dx)
        dp = NumberShiftableVector.unwrap(dp)
        new_state = self.method.advance_state(self.state, dx, dp, ddep)
        return TimesteppablePicState(self.method, new_state)

    __radd__ = __add__


class FieldToParticleRhsCalculator:
    """cloud.py:238-288: forces from E and B = mu H, ZeroVector where absent."""

    def __init__(self, method, maxwell_op):
        self.method = method
        self.maxwell_op = maxwell_op

    def __call__(self, t, fields_f, state_f):
        fields = fields_f()
        state = state_f()
        e, h = self.maxwell_op.split_eh(fields)
        idx = 0
        e_arg = []
        for use in self.maxwell_op.get_eh_subset()[0:3]:
            if use:
                e_arg.append(e[idx])
                idx += 1
            else:
                e_arg.append(_internal.ZeroVector())
        idx = 0
        b_arg = []
        for use in self.maxwell_op.get_eh_subset()[3:6]:
            if use:
                b_arg.append(self.maxwell_op.mu * h[idx])
                idx += 1
            else:
                b_arg.append(_internal.ZeroVector())
        velocities = self.method.velocities(state)
        forces = self.method.pusher.forces(state, velocities, *(e_arg + b_arg))
        return make_obj_array([
            0, NumberShiftableVector(forces, signaller=state.particle_number_shift_signaller),
            0])


class ParticleRhsCalculator:
    """cloud.py:293-311: dx/dt = v."""

    def __init__(self, method, maxwell_op=None):
        self.method = method
        self.maxwell_op = maxwell_op

    def __call__(self, t, fields_f, state_f):
        state = state_f()
        velocities = self.method.velocities(state)
        return make_obj_array([
            NumberShiftableVector(velocities, signaller=state.particle_number_shift_signaller),
            0, self.method.depositor.rhs(state)])


class ParticleToFieldRhsCalculator:
    """cloud.py:225-233: the particles' contribution to the field RHS,
    dE/dt = -j / epsilon, with j deposited by the method's depositor (one
    device array per component)."""

    def __init__(self, method, maxwell_op):
        self.method = method
        self.maxwell_op = maxwell_op

    def __call__(self, t, fields_f, state_f):
        j = self.method.deposit_j(state_f())
        return self.maxwell_op.assemble_fields(e=[-1 / self.maxwell_op.epsilon * jc for jc in j])


class PicMethod:
    """cloud.py:316-603."""

    def __init__(self, discr, units, depositor, pusher, dimensions_pos, dimensions_velocity,
                 debug=set()):
        self.units = units
        self.discretization = discr
        self.debug = debug
        self.depositor = depositor
        self.pusher = pusher
        self.dimensions_mesh = discr.dimensions
        self.dimensions_pos = dimensions_pos
        self.dimensions_velocity = dimensions_velocity
        if not (dimensions_pos == dimensions_velocity == discr.dimensions):
            raise ValueError("positions, velocities and mesh must share the dimension "
                             "(2d2v or 3d3v)")
        self.mesh_data = _internal.MeshData(discr.dimensions)
        self.mesh_data.fill_from_hedge(discr)
        self.depositor.initialize(self)
        self.pusher.initialize(self)
        self.find_counters = _internal.FindEventCounters()

    def make_state(self):
        state = PicState(self)
        state.particle_number_shift_signaller.subscribe_with_state(self.depositor)
        state.particle_number_shift_signaller.subscribe_with_state(self.pusher)
        return state

    def get_dimensionality_suffix(self):
        return "%dd%dv" % (self.dimensions_pos, self.dimensions_velocity)

    def get_shape_function_class(self):
        return _internal.PolynomialShapeFunction

    def velocities(self, state):
        try:
            return state.derived_quantity_cache["velocities"]
        except KeyError:
            result = _internal.get_velocities(state.particle_state,
                                              self.units.VACUUM_LIGHT_SPEED())
            state.derived_quantity_cache["velocities"] = result
            return result

    def mean_beta(self, state):
        if len(state):
            return (self.velocities(state).mean(dim=0) / self.units.VACUUM_LIGHT_SPEED())
        return torch.zeros(self.dimensions_velocity, dtype=torch.float64, device="cuda")

    def add_particles(self, state, iterable, maxcount=None):
        """cloud.py:441-490: tuples (position, velocity, charge, mass)."""
        items = []
        for pos, vel, charge, mass in iterable:
            if maxcount is not None:
                if maxcount == 0:
                    break
                maxcount -= 1
            assert len(pos) == self.dimensions_pos
            assert len(vel) == self.dimensions_velocity
            vel = np.asarray(vel, dtype=np.float64)
            mom = mass * self.units.gamma_from_v(vel) * vel
            items.append((np.asarray(pos, dtype=np.float64), mom, charge, mass))
        if not items:
            return
        self.add_particle_arrays(state, np.array([i[0] for i in items]),
                                 np.array([i[1] for i in items]),
                                 np.array([i[2] for i in items], dtype=np.float64),
                                 np.array([i[3] for i in items], dtype=np.float64))

    def add_particle_arrays(self, state, positions, momenta, charges, masses):
        """Vectorized add_particles; particles outside the mesh are skipped."""
        ps = state.particle_state
        x = torch.as_tensor(np.ascontiguousarray(positions, dtype=np.float64), device="cuda")
        el = self.mesh_data.find_containing_elements(x)
        keep = el != np.int32(np.uint32(_internal.INVALID_ELEMENT).view(np.int32))
        n_add = int(keep.sum())
        if n_add == 0:
            return
        n0 = ps.particle_count
        if n0 + n_add > ps.positions.shape[0]:
            state.resize(max(128, 2 * (n0 + n_add)))
        sel = keep.nonzero().squeeze(1)
        ps.containing_elements[n0:n0 + n_add] = el[sel]
        ps.positions[n0:n0 + n_add] = x[sel]
        ps.momenta[n0:n0 + n_add] = torch.as_tensor(
            np.ascontiguousarray(momenta, dtype=np.float64), device="cuda")[sel]
        ps.charges[n0:n0 + n_add] = torch.as_tensor(np.asarray(charges, dtype=np.float64),
                                                    device="cuda")[sel]
        ps.masses[n0:n0 + n_add] = torch.as_tensor(np.asarray(masses, dtype=np.float64),
                                                   device="cuda")[sel]
        ps.particle_count = n0 + n_add
        self.check_containment(state)
        state.particle_number_shift_signaller.note_change_size(ps.particle_count)
        state.derived_quantity_cache.clear()

    def check_containment(self, state):
        inv = int(np.uint32(_internal.INVALID_ELEMENT).view(np.int32))
        assert bool((state.containing_elements != inv).all())

    def upkeep(self, state):
        self.depositor.upkeep(state)
        self.pusher.upkeep(state)
        state.vis_listener.clear()

    # deposition ----------------------------------------------------------
    def deposit_densities(self, state):
        """(rho, j) with j as vdim x nodes (cloud.py:513-526)."""
        def getter():
            rho, j = self.depositor.deposit_densites(state, self.velocities(state))
            return rho, j.T.contiguous()
        if "rho" in state.derived_quantity_cache and "j" in state.derived_quantity_cache:
            return state.derived_quantity_cache["rho"], state.derived_quantity_cache["j"]
        rho, j = getter()
        state.derived_quantity_cache["rho"] = rho
        state.derived_quantity_cache["j"] = j
        return rho, j

    def deposit_j(self, state):
        """j as vdim x nodes (cloud.py:528-540)."""
        return state.get_derived_quantity_from_cache(
            "j", lambda: self.depositor.deposit_j(state, self.velocities(state)).T.contiguous())

    def deposit_rho(self, state):
        return state.get_derived_quantity_from_cache(
            "rho", lambda: self.depositor.deposit_rho(state))

    # time advance ----------------------------------------------------------
    def advance_state(self, state, dx, dp, ddep):
        """cloud.py:554-603: x += dx, p += dp, relocalize, kill."""
        ps = state.particle_state
        cnt = ps.particle_count
        positions = ps.positions.clone()
        momenta = ps.momenta.clone()
        positions[:cnt] += dx
        momenta[:cnt] += dp
        new_state = PicState(
            self, particle_count=cnt, containing_elements=ps.containing_elements,
            positions=positions, momenta=momenta, charges=ps.charges, masses=ps.masses,
            depositor_state=self.depositor.advance_state(state, ddep),
            pusher_state=self.pusher.advance_state(state),
            pnss=state.particle_number_shift_signaller, vis_listener=state.vis_listener)
        _internal.update_containing_elements(
            self.mesh_data, new_state.particle_state,
            new_state.particle_number_shift_signaller, self.find_counters)
        return new_state
