"""Low-storage fourth-order Runge-Kutta, as hedge's LSRK4TimeStepper provides it
to the reference (used at test/static_field.py:265,317,409 and
test/test_pyrticle.py:162-174).

Five stages, 2N storage: per stage ``residual = a*residual + dt*rhs(t+c*dt, y)``
then ``y = y + b*residual``; the residual persists between calls and starts as
``0*rhs(t, y)``.  Coefficients: M. H. Carpenter and C. A. Kennedy, "Fourth-order
2N-storage Runge-Kutta schemes", NASA TM-109112 (1994), solution 3.
"""

from __future__ import annotations

import numpy as np

RK4A = [0.0,
        -567301805773 / 1357537059087,
        -2404267990393 / 2016746695238,
        -3550918686646 / 2091501179385,
        -1275806237668 / 842570457699]
RK4B = [1432997174477 / 9575080441755,
        5161836677717 / 13612068292357,
        1720146321549 / 2090206949498,
        3134564353537 / 4481467310338,
        2277821191437 / 14882151754819]
RK4C = [0.0,
        1432997174477 / 9575080441755,
        2526269341429 / 6820363183101,
        2006345519317 / 3224310063776,
        2802321613138 / 2924317926251]


class LSRK4TimeStepper:
    dt_fudge = 1

    def __init__(self):
        self.coeffs = np.array([RK4A, RK4B, RK4C]).T

    def __call__(self, y, t, dt, rhs):
        try:
            self.residual
        except AttributeError:
            self.residual = 0 * rhs(t, y)
        for a, b, c in self.coeffs:
            this_rhs = rhs(t + c * dt, y)
            self.residual = a * self.residual + dt * this_rhs
            del this_rhs
            y = y + b * self.residual
        return y
