"""Fused LSRK4 particle push in static fields.

The reference advances particles through hedge's LSRK4 stepper with Python-
level RHS calls per stage (cloud.py:159-311).  For static E and B the whole
step -- five stages of v(p), interpolated q(E + v x B), residual updates and
relocalization with kills -- runs here in one kernel per step
(pyr_lsrk4_steps), in the same arithmetic as the per-stage path, so both give
bit-identical states.  Kills follow the reference's per-stage swap-remove order;
subscribers of the state's number-shift signaller receive the net moves.
"""

from __future__ import annotations

import torch

from . import device as dv
from .cloud import PicState, TimesteppablePicState


class FusedParticleStepper:
    """``stepper(ts_state, t, dt, nsteps=1)`` for the particle RHS [v, F] in
    static fields (E nodal, B = mu H nodal; None for zero components)."""

    def __init__(self, method, e_fields, b_fields, precision="fp64"):
        from .pusher import AverageParticlePusher, MonomialParticlePusher
        self.average = isinstance(method.pusher, AverageParticlePusher)
        if not (self.average or isinstance(method.pusher, MonomialParticlePusher)):
            raise ValueError("the fused step supports MonomialParticlePusher and "
                             "AverageParticlePusher; step other pushers through the "
                             "per-stage RHS")
        if self.average and precision != "fp64":
            raise ValueError("the averaging pusher's step is FP64 only")
        self.method = method
        self.precision = dv.FP64 if precision == "fp64" else dv.FP32C
        fields = [None if f is None else f.contiguous() for f in list(e_fields) + list(b_fields)]
        assert len(fields) == 6
        self._fields = fields
        method.mesh_data.device.set_static_fields(fields)
        self.counters = torch.zeros(4, dtype=torch.int64, device="cuda")

    def __call__(self, ts_state, t, dt, nsteps=1):
        state = ts_state.state
        ps = state.particle_state
        n = ps.particle_count
        ids = torch.arange(n, dtype=torch.int64, device="cuda")
        dev = self.method.mesh_data.device
        c = self.method.units.VACUUM_LIGHT_SPEED()
        if self.average:
            # the depositor's shape function and element reach, as the
            # averaging pusher's per-stage forces use them (statistics and vis
            # vectors are left to the per-stage path)
            dep = self.method.depositor.backend
            sf = dep.shape_function
            if sf is None:
                raise RuntimeError("shape function never set")
            n_new = dev.lsrk4_steps_average(
                nsteps, dt, c, sf.radius, sf.exponent, sf.normalizer,
                self.method.pusher.backend.integral_weights, n, ps.positions[:n],
                ps.momenta[:n], ps.charges[:n], ps.masses[:n], ps.containing_elements[:n],
                ids=ids, counters=self.counters)
        else:
            n_new = dev.lsrk4_steps(
                nsteps, dt, c, n, ps.positions[:n], ps.momenta[:n], ps.charges[:n],
                ps.masses[:n], ps.containing_elements[:n], ids=ids, counters=self.counters,
                precision=self.precision)
        ps.particle_count = n_new
        sig = state.particle_number_shift_signaller
        moved = (ids[:n_new] != torch.arange(n_new, device="cuda")).nonzero().squeeze(1)
        if moved.numel():
            sig.note_move(ids[moved].to(torch.int32), moved.to(torch.int32), 1)
        if n_new != n:
            sig.note_change_size(n_new)
        new_state = PicState(
            self.method, particle_count=n_new, containing_elements=ps.containing_elements,
            positions=ps.positions, momenta=ps.momenta, charges=ps.charges, masses=ps.masses,
            depositor_state=state.depositor_state, pusher_state=state.pusher_state, pnss=sig,
            vis_listener=state.vis_listener)
        return TimesteppablePicState(self.method, new_state)
