"""A minimal stand-in for the Maxwell operator interface the particle RHS
calculators read (cloud.py:238-288): ``split_eh``, ``get_eh_subset``, ``mu``,
``epsilon``, ``assemble_eh`` and ``assemble_fields``.  The field solve itself
is outside this package; fields are nodal device vectors of length K*Np.
"""

from __future__ import annotations

import numpy as np

from .units import SIUnitsWithNaturalConstants as _U


class StaticFieldOperator:
    """Field container for static E and H.

    ``eh_subset`` is the hedge-style 6-tuple of booleans selecting which of
    (ex, ey, ez, hx, hy, hz) are carried, e.g. TE in 2D: (True, True, False,
    False, False, True).
    """

    def __init__(self, eh_subset=(True,) * 6, epsilon=_U.EPSILON0, mu=_U.MU0):
        self.eh_subset = tuple(bool(b) for b in eh_subset)
        self.epsilon = epsilon
        self.mu = mu

    def get_eh_subset(self):
        return self.eh_subset

    def count_subset(self, which):
        return sum(1 for b in which if b)

    def assemble_eh(self, e=None, h=None):
        ne = self.count_subset(self.eh_subset[:3])
        nh = self.count_subset(self.eh_subset[3:])
        e = list(e) if e is not None else []
        h = list(h) if h is not None else []
        assert len(e) == ne and len(h) == nh
        out = np.empty(ne + nh, dtype=object)
        for i, v in enumerate(e + h):
            out[i] = v
        return out

    def assemble_fields(self, e=None, h=None):
        """Like assemble_eh, with an absent part filled with zeros (scalar 0
        entries, which add to nodal vectors as zero fields)."""
        ne = self.count_subset(self.eh_subset[:3])
        nh = self.count_subset(self.eh_subset[3:])
        e = [0] * ne if e is None else list(e)
        h = [0] * nh if h is None else list(h)
        return self.assemble_eh(e, h)

    def split_eh(self, fields):
        ne = self.count_subset(self.eh_subset[:3])
        return list(fields[:ne]), list(fields[ne:])
