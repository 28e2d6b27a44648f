"""Number-shift signalling, visualization storage and warning forwarding
(pyrticle/tools.py:31-304), over device tensors.

Kills are swap-removes (particle_state.hpp:361-407); vectors indexed by
particle number (RK residuals, visualization data) must follow them.  Moves
arrive batched: ``note_move(orig, dest, size)`` takes either two ints or two
int32 index tensors of equal length, applied as ``v[dest] = v[orig]`` (the
sources lie past the new count and the destinations below it, so the batch has
no ordering hazards).
"""

from __future__ import annotations

import warnings
from weakref import WeakKeyDictionary

import numpy as np
import torch

from . import _internal
from ._internal import PolynomialShapeFunction  # noqa: F401  (tools.py:263 alias)


class WarningForwarder(_internal.WarningListener):
    def note_warning(self, message, filename, lineno):
        warnings.warn_explicit(message, UserWarning, filename, lineno)


class WarningIgnorer(_internal.WarningListener):
    def note_warning(self, message, filename, lineno):
        pass


warning_forwarder = WarningForwarder()


def _as_index(v, device):
    if isinstance(v, torch.Tensor):
        return v.to(device=device, dtype=torch.long)
    return torch.tensor([int(v)], dtype=torch.long, device=device)


class NumberShiftMultiplexer(_internal.NumberShiftListener):
    def __init__(self, name=None):
        self.subscribers = WeakKeyDictionary()
        self.name = name

    def subscribe(self, subscriber):
        self.subscribers[subscriber] = None

    def note_change_size(self, new_size):
        for subscriber in list(self.subscribers.keys()):
            subscriber.note_change_size(new_size)

    def note_move(self, orig, dest, size):
        for subscriber in list(self.subscribers.keys()):
            subscriber.note_move(orig, dest, size)

    def note_reset(self, start, size):
        for subscriber in list(self.subscribers.keys()):
            subscriber.note_reset(start, size)


class StatePassingNumberShiftMultiplexer(NumberShiftMultiplexer):
    def __init__(self, state):
        NumberShiftMultiplexer.__init__(self)
        self.subscribers_with_state = WeakKeyDictionary()
        self.state = state

    def subscribe_with_state(self, subscriber):
        self.subscribers_with_state[subscriber] = None

    def note_change_size(self, new_size):
        NumberShiftMultiplexer.note_change_size(self, new_size)
        for subscriber in list(self.subscribers_with_state.keys()):
            subscriber.note_change_size(self.state, new_size)

    def note_move(self, orig, dest, size):
        NumberShiftMultiplexer.note_move(self, orig, dest, size)
        for subscriber in list(self.subscribers_with_state.keys()):
            subscriber.note_move(self.state, orig, dest, size)

    def note_reset(self, start, size):
        NumberShiftMultiplexer.note_reset(self, start, size)
        for subscriber in list(self.subscribers_with_state.keys()):
            subscriber.note_reset(self.state, start, size)


class NumberShiftableVector(_internal.NumberShiftListener):
    """A per-particle device vector that follows kills (tools.py:106-178)."""

    def __init__(self, vector, signaller):
        self.vector = vector
        self.signaller = signaller
        signaller.subscribe(self)

    @staticmethod
    def unwrap(instance):
        if isinstance(instance, NumberShiftableVector):
            return instance.vector
        return instance

    def __len__(self):
        return len(self.vector)

    def __add__(self, other):
        return NumberShiftableVector(self.vector + self.unwrap(other), self.signaller)

    __radd__ = __add__

    def __iadd__(self, other):
        self.vector += self.unwrap(other)
        return self

    def __mul__(self, other):
        return NumberShiftableVector(self.vector * self.unwrap(other), self.signaller)

    __rmul__ = __mul__

    def note_change_size(self, new_size):
        old = self.vector.shape[0]
        if new_size > old:
            grown = torch.empty((new_size,) + tuple(self.vector.shape[1:]),
                                dtype=self.vector.dtype, device=self.vector.device)
            grown[:old] = self.vector
            self.vector = grown
        elif new_size < old:
            self.vector = self.vector[:new_size]

    def note_move(self, orig, dest, size):
        dev = self.vector.device
        self.vector[_as_index(dest, dev)] = self.vector[_as_index(orig, dev)]

    def note_reset(self, start, size):
        self.vector[start:start + size] = 0


class MapStorageVisualizationListener(_internal.VisualizationListener):
    """tools.py:287-304: keeps the latest particle and mesh vis vectors."""

    def __init__(self, particle_number_shift_signaller):
        self.mesh_vis_map = {}
        self.particle_vis_map = {}
        self.particle_number_shift_signaller = particle_number_shift_signaller

    def store_mesh_vis_vector(self, name, vec):
        self.mesh_vis_map[name] = vec

    def store_particle_vis_vector(self, name, vec):
        self.particle_vis_map[name] = NumberShiftableVector(
            vec, self.particle_number_shift_signaller)

    def clear(self):
        self.mesh_vis_map.clear()
        self.particle_vis_map.clear()


def make_obj_array(entries):
    out = np.empty(len(entries), dtype=object)
    for i, e in enumerate(entries):
        out[i] = e
    return out


def is_zero(x):
    if isinstance(x, (int, float)):
        return x == 0
    if isinstance(x, np.ndarray) and x.dtype == object:
        return all(is_zero(e) for e in x)
    return False
